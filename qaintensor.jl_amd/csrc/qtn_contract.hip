// Axis permutation and the pairwise contraction step (permute + complex
// GEMM + permute), i.e. what one pairwise step of TensorOperations.ncon does
// for Qaintensor's contract() (src/contract.jl:242-264).
//
// Permute: when the fastest axis is kept, a gather with unit-stride runs on
// both sides.  Otherwise the fastest input axis and the axis that becomes
// the fastest output axis are tiled 32x32 through LDS so that both the read
// and the write are coalesced; all other axes are enumerated by blockIdx.z.
#include <algorithm>
#include <numeric>

#include <mutex>

#include "qtn_internal.hpp"

namespace qtn {
namespace {

constexpr int kMaxNd = 16;

struct PermDesc {
  int nd;
  int64_t odims[kMaxNd];     // output dims
  int64_t istride_o[kMaxNd]; // input stride of output axis d
};

// Gather permute: one thread per output element (used when the fastest
// axis is unchanged, so both sides are unit-stride along axis 0).
__global__ void permute_gather(const double2* __restrict__ in,
                               double2* __restrict__ out, int64_t total,
                               PermDesc d) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; idx < total; idx += stride) {
    int64_t rem = idx, off = 0;
    for (int a = 0; a < d.nd; ++a) {
      const int64_t q = rem / d.odims[a];
      off += (rem - q * d.odims[a]) * d.istride_o[a];
      rem = q;
    }
    out[idx] = in[off];
  }
}

struct TileDesc {
  int64_t dA;          // size of input axis 0 (an output axis q != 0)
  int64_t dB;          // size of output axis 0 (input axis p != 0)
  int64_t in_sB;       // input stride of axis p
  int64_t out_sA;      // output stride of output axis q
  int nrest;           // remaining axes
  int64_t rdims[kMaxNd];
  int64_t rin[kMaxNd];
  int64_t rout[kMaxNd];
};

__global__ __launch_bounds__(256) void permute_tiled(
    const double2* __restrict__ in, double2* __restrict__ out, TileDesc d) {
  __shared__ double2 tile[32][33];
  const int tx = threadIdx.x & 31;  // 0..31
  const int ty = threadIdx.x >> 5;  // 0..7
  const int64_t a0 = (int64_t)blockIdx.x * 32;
  const int64_t b0 = (int64_t)blockIdx.y * 32;
  // decode the rest index
  int64_t rem = blockIdx.z, in_off = 0, out_off = 0;
  for (int r = 0; r < d.nrest; ++r) {
    const int64_t q = rem / d.rdims[r];
    const int64_t i = rem - q * d.rdims[r];
    in_off += i * d.rin[r];
    out_off += i * d.rout[r];
    rem = q;
  }
  // read: a fastest (input axis 0 is unit stride)
  for (int j = ty; j < 32; j += 8) {
    const int64_t a = a0 + tx, b = b0 + j;
    if (a < d.dA && b < d.dB) tile[j][tx] = in[in_off + a + b * d.in_sB];
  }
  __syncthreads();
  // write: b fastest (output axis 0 is unit stride)
  for (int j = ty; j < 32; j += 8) {
    const int64_t b = b0 + tx, a = a0 + j;
    if (a < d.dA && b < d.dB) out[out_off + b + a * d.out_sA] = tile[tx][j];
  }
}

}  // namespace

int32_t launch_permute(hipStream_t s, const double2* in, int nd,
                       const int64_t* dims, const int* perm0, double2* out) {
  if (nd < 0 || nd > kMaxNd) {
    set_error("permute: rank must be between 0 and 16");
    return QTN_ERR_ARG;
  }
  int64_t total = 1;
  for (int i = 0; i < nd; ++i) total *= dims[i];
  if (total == 0) return QTN_OK;
  int64_t istride[kMaxNd];
  int64_t acc = 1;
  for (int i = 0; i < nd; ++i) {
    istride[i] = acc;
    acc *= dims[i];
  }
  // Drop unit axes: they change neither side's layout.
  bool identity = true;
  {
    int last = -1;
    for (int d = 0; d < nd; ++d) {
      if (dims[perm0[d]] == 1) continue;
      if (perm0[d] < last) identity = false;
      last = perm0[d];
    }
  }
  if (identity) {
    QTN_CHECK_HIP(hipMemcpyAsync(out, in, total * sizeof(double2),
                                 hipMemcpyDeviceToDevice, s));
    return QTN_OK;
  }
  int64_t ostride[kMaxNd];
  acc = 1;
  for (int d = 0; d < nd; ++d) {
    ostride[d] = acc;
    acc *= dims[perm0[d]];
  }
  // First non-unit input axis and first non-unit output axis.
  int in_fast = -1;
  for (int i = 0; i < nd; ++i)
    if (dims[i] != 1) {
      in_fast = i;
      break;
    }
  int out_fast = -1;
  for (int d = 0; d < nd; ++d)
    if (dims[perm0[d]] != 1) {
      out_fast = d;
      break;
    }
  const bool fast_kept = (perm0[out_fast] == in_fast);
  if (fast_kept || dims[in_fast] < 4 || dims[perm0[out_fast]] < 4) {
    PermDesc pd;
    pd.nd = 0;
    for (int d = 0; d < nd; ++d) {
      if (dims[perm0[d]] == 1) continue;
      pd.odims[pd.nd] = dims[perm0[d]];
      pd.istride_o[pd.nd] = istride[perm0[d]];
      ++pd.nd;
    }
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(permute_gather, dim3((unsigned)blocks), dim3(256), 0,
                       s, in, out, total, pd);
    QTN_CHECK_HIP(hipGetLastError());
    return QTN_OK;
  }
  TileDesc td;
  const int p = perm0[out_fast];  // input axis that becomes output-fastest
  int q = -1;                     // output axis that holds input axis in_fast
  for (int d = 0; d < nd; ++d)
    if (perm0[d] == in_fast) q = d;
  td.dA = dims[in_fast];
  td.dB = dims[p];
  td.in_sB = istride[p];
  td.out_sA = ostride[q];
  td.nrest = 0;
  int64_t nrest_total = 1;
  for (int d = 0; d < nd; ++d) {
    const int ax = perm0[d];
    if (ax == p || ax == in_fast || dims[ax] == 1) continue;
    td.rdims[td.nrest] = dims[ax];
    td.rin[td.nrest] = istride[ax];
    td.rout[td.nrest] = ostride[d];
    nrest_total *= dims[ax];
    ++td.nrest;
  }
  const int64_t ga = (td.dA + 31) / 32, gb = (td.dB + 31) / 32;
  if (ga > INT32_MAX || gb > 65535 || nrest_total > 65535) {
    // Fall back to the gather form for extreme shapes.
    PermDesc pd;
    pd.nd = 0;
    for (int d = 0; d < nd; ++d) {
      if (dims[perm0[d]] == 1) continue;
      pd.odims[pd.nd] = dims[perm0[d]];
      pd.istride_o[pd.nd] = istride[perm0[d]];
      ++pd.nd;
    }
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(permute_gather, dim3((unsigned)blocks), dim3(256), 0,
                       s, in, out, total, pd);
    QTN_CHECK_HIP(hipGetLastError());
    return QTN_OK;
  }
  hipLaunchKernelGGL(permute_tiled,
                     dim3((unsigned)ga, (unsigned)gb, (unsigned)nrest_total),
                     dim3(256), 0, s, in, out, td);
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}

namespace {

// Host-side plan of one pairwise contraction (fixed-size arrays: planning
// allocates nothing).
struct PairPlan {
  int permA[kMaxNd], permB[kMaxNd];  // 0-based
  int ndA = 0, ndB = 0;
  int64_t M = 1, N = 1, K = 1;
  bool transA = false, transB = false;  // use op 'T' instead of permuting
  bool permuteA = false, permuteB = false;
  int64_t natdims[2 * kMaxNd];        // natural output dims
  int outperm0[2 * kMaxNd];           // 0-based
  int ndC = 0;
  bool permuteC = false;
};

bool is_identity(const int* p, int n) {
  for (int i = 0; i < n; ++i)
    if (p[i] != i) return false;
  return true;
}

int32_t make_plan(int ndA, const int64_t* dimsA, int ndB,
                  const int64_t* dimsB, int nc, const int32_t* legsA,
                  const int32_t* legsB, const int32_t* outperm,
                  PairPlan& pl) {
  QTN_ARG(ndA >= 0 && ndA <= kMaxNd && ndB >= 0 && ndB <= kMaxNd,
          "contract_pair: tensor rank must be between 0 and 16");
  QTN_ARG(nc >= 0 && nc <= ndA && nc <= ndB,
          "contract_pair: bad number of contracted legs");
  for (int i = 0; i < ndA; ++i)
    QTN_ARG(dimsA[i] >= 0, "contract_pair: negative dimension in A");
  for (int i = 0; i < ndB; ++i)
    QTN_ARG(dimsB[i] >= 0, "contract_pair: negative dimension in B");
  char usedA[kMaxNd] = {}, usedB[kMaxNd] = {};
  int la[kMaxNd], lb[kMaxNd], fa[kMaxNd], fb[kMaxNd];
  for (int i = 0; i < nc; ++i) {
    const int a = legsA[i] - 1, b = legsB[i] - 1;
    QTN_ARG(a >= 0 && a < ndA && b >= 0 && b < ndB,
            "contract_pair: leg number out of range");
    QTN_ARG(!usedA[a] && !usedB[b], "contract_pair: leg contracted twice");
    QTN_ARG(dimsA[a] == dimsB[b],
            "Dimensions of contraction legs do not match");
    usedA[a] = usedB[b] = 1;
    la[i] = a;
    lb[i] = b;
  }
  int nfa = 0, nfb = 0;
  for (int i = 0; i < ndA; ++i)
    if (!usedA[i]) fa[nfa++] = i;
  for (int i = 0; i < ndB; ++i)
    if (!usedB[i]) fb[nfb++] = i;
  pl.ndA = ndA;
  pl.ndB = ndB;
  pl.M = pl.N = pl.K = 1;
  for (int i = 0; i < nfa; ++i) pl.M *= dimsA[fa[i]];
  for (int i = 0; i < nfb; ++i) pl.N *= dimsB[fb[i]];
  for (int i = 0; i < nc; ++i) pl.K *= dimsA[la[i]];
  // A as (free x contracted): permA = fa ++ la; or (contracted x free).
  int altA[kMaxNd];
  for (int i = 0; i < nfa; ++i) pl.permA[i] = fa[i];
  for (int i = 0; i < nc; ++i) pl.permA[nfa + i] = la[i];
  for (int i = 0; i < nc; ++i) altA[i] = la[i];
  for (int i = 0; i < nfa; ++i) altA[nc + i] = fa[i];
  if (is_identity(pl.permA, ndA)) {
    pl.transA = false;
    pl.permuteA = false;
  } else if (is_identity(altA, ndA)) {
    pl.transA = true;
    pl.permuteA = false;
  } else {
    pl.transA = false;
    pl.permuteA = true;
  }
  int altB[kMaxNd];
  for (int i = 0; i < nc; ++i) pl.permB[i] = lb[i];
  for (int i = 0; i < nfb; ++i) pl.permB[nc + i] = fb[i];
  for (int i = 0; i < nfb; ++i) altB[i] = fb[i];
  for (int i = 0; i < nc; ++i) altB[nfb + i] = lb[i];
  if (is_identity(pl.permB, ndB)) {
    pl.transB = false;
    pl.permuteB = false;
  } else if (is_identity(altB, ndB)) {
    pl.transB = true;
    pl.permuteB = false;
  } else {
    pl.transB = false;
    pl.permuteB = true;
  }
  pl.ndC = 0;
  for (int i = 0; i < nfa; ++i) pl.natdims[pl.ndC++] = dimsA[fa[i]];
  for (int i = 0; i < nfb; ++i) pl.natdims[pl.ndC++] = dimsB[fb[i]];
  char seen[2 * kMaxNd] = {};
  for (int i = 0; i < pl.ndC; ++i) {
    const int p = outperm ? outperm[i] - 1 : i;
    QTN_ARG(p >= 0 && p < pl.ndC && !seen[p],
            "contract_pair: outperm is not a permutation");
    seen[p] = 1;
    pl.outperm0[i] = p;
  }
  pl.permuteC = !is_identity(pl.outperm0, pl.ndC);
  QTN_ARG(!pl.permuteC || pl.ndC <= kMaxNd,
          "contract_pair: a permuted result has at most 16 legs");
  return QTN_OK;
}

// Element counts of the plan's scratch: the permuted A and B and the
// product in natural leg order (0 where that step is not needed).
void plan_scratch(const PairPlan& pl, int64_t* nA, int64_t* nB,
                  int64_t* nC) {
  *nA = pl.permuteA ? std::max<int64_t>(1, pl.M * pl.K) : 0;
  *nB = pl.permuteB ? std::max<int64_t>(1, pl.K * pl.N) : 0;
  *nC = pl.permuteC ? std::max<int64_t>(1, pl.M * pl.N) : 0;
}

// Permutes and GEMM of one pairwise contraction on stream s, with the
// scratch given (sA, sB, sC of plan_scratch's sizes; unused ones may be
// null).  Launches only; allocates nothing.
int32_t run_plan_with(hipStream_t s, const PairPlan& pl, const double2* A,
                      const int64_t* dimsA, const double2* B,
                      const int64_t* dimsB, double2* sA, double2* sB,
                      double2* sC, double2* C) {
  const double2* Am = A;
  const double2* Bm = B;
  if (pl.permuteA) {
    QTN_TRY(launch_permute(s, A, pl.ndA, dimsA, pl.permA, sA));
    Am = sA;
  }
  if (pl.permuteB) {
    QTN_TRY(launch_permute(s, B, pl.ndB, dimsB, pl.permB, sB));
    Bm = sB;
  }
  double2* Cn = pl.permuteC ? sC : C;
  const int opA = pl.transA ? 1 : 0;
  const int opB = pl.transB ? 1 : 0;
  const int64_t lda = pl.transA ? std::max<int64_t>(1, pl.K)
                                : std::max<int64_t>(1, pl.M);
  const int64_t ldb = pl.transB ? std::max<int64_t>(1, pl.N)
                                : std::max<int64_t>(1, pl.K);
  QTN_TRY(launch_zgemm(s, opA, opB, pl.M, pl.N, pl.K, make_double2(1, 0), Am,
                       lda, 0, Bm, ldb, 0, make_double2(0, 0), Cn,
                       std::max<int64_t>(1, pl.M), 0, 1));
  if (pl.permuteC)
    QTN_TRY(launch_permute(s, Cn, pl.ndC, pl.natdims, pl.outperm0, C));
  return QTN_OK;
}

// The plan with the context's scratch slots 0..2.
int32_t run_plan(qtn_ctx* ctx, const PairPlan& pl, const double2* A,
                 const int64_t* dimsA, const double2* B,
                 const int64_t* dimsB, double2* C) {
  int64_t nA, nB, nC;
  plan_scratch(pl, &nA, &nB, &nC);
  void *pA = nullptr, *pB = nullptr, *pC = nullptr;
  if (nA) QTN_TRY(ctx->buffer(0, nA * sizeof(double2), &pA));
  if (nB) QTN_TRY(ctx->buffer(1, nB * sizeof(double2), &pB));
  if (nC) QTN_TRY(ctx->buffer(2, nC * sizeof(double2), &pC));
  return run_plan_with(ctx->stream, pl, A, dimsA, B, dimsB, (double2*)pA,
                       (double2*)pB, (double2*)pC, C);
}

}  // namespace
}  // namespace qtn

// ---- the allocation-free pairwise contraction -----------------------------
// scratch: the device memory qtn_contract_pair_dev_noalloc takes its scratch
// from (the permuted operands and the product in natural leg order), 32 MiB
// per GPU for the life of the process.  Host code never passes the array's
// name as an address (on the host it is only a stand-in): each GPU's copy
// sits at its own device address, looked up once per device by
// noalloc_scratch_init (qtn_ctx_create calls it with the device current)
// and kept in g_scratch_addr.
constexpr int64_t kNoAllocScratchBytes = (int64_t)32 << 20;
constexpr int64_t kNoAllocScratchElems = kNoAllocScratchBytes / 16;
static __device__ __align__(256) double2 scratch[kNoAllocScratchElems];

namespace qtn {
namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxScratchDevices] = {};
}  // namespace

int32_t noalloc_scratch_init(int device) {
  QTN_ARG(device >= 0 && device < kMaxScratchDevices,
          "qtn_ctx_create: device number too large for the no-allocation "
          "scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_addr[device]) return QTN_OK;
  void* p = nullptr;
  QTN_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_addr[device] = p;
  return QTN_OK;
}
}  // namespace qtn

extern "C" {

int32_t qtn_contract_pair_dev_noalloc(qtn_ctx* ctx, const double* A,
                                      int32_t ndA, const int64_t* dimsA,
                                      const double* B, int32_t ndB,
                                      const int64_t* dimsB, int32_t nc,
                                      const int32_t* legsA,
                                      const int32_t* legsB,
                                      const int32_t* outperm,
                                      double* C) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr, "qtn_contract_pair_dev_noalloc: ctx is NULL");
  QTN_ARG((ndA == 0 || dimsA) && (ndB == 0 || dimsB) &&
              (nc == 0 || (legsA && legsB)),
          "qtn_contract_pair_dev_noalloc: NULL argument");
  qtn::PairPlan pl;
  QTN_TRY(qtn::make_plan(ndA, dimsA, ndB, dimsB, nc, legsA, legsB, outperm,
                         pl));
  int64_t nA, nB, nC;
  qtn::plan_scratch(pl, &nA, &nB, &nC);
  // each piece starts 256-byte aligned (16 complex doubles)
  const int64_t rA = (nA + 15) / 16 * 16, rB = (nB + 15) / 16 * 16,
                rC = (nC + 15) / 16 * 16;
  QTN_ARG(rA + rB + rC <= kNoAllocScratchElems,
          "qtn_contract_pair_dev_noalloc: the contraction needs more "
          "scratch than the 32 MiB this entry point has (permuted operands "
          "and product); use qtn_contract_pair_dev");
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(qtn::g_scratch_mu);
    base = qtn::g_scratch_addr[ctx->device];
  }
  QTN_ARG(base != nullptr, "qtn_contract_pair_dev_noalloc: no scratch for "
                           "this device (context not created by "
                           "qtn_ctx_create)");
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  double2* sA = (double2*)base;
  double2* sB = sA + rA;
  double2* sC = sB + rB;
  return qtn::run_plan_with(ctx->stream, pl, (const double2*)A, dimsA,
                            (const double2*)B, dimsB, nA ? sA : nullptr,
                            nB ? sB : nullptr, nC ? sC : nullptr,
                            (double2*)C);
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_contract_pair_dev(qtn_ctx* ctx, const double* A, int32_t ndA,
                              const int64_t* dimsA, const double* B,
                              int32_t ndB, const int64_t* dimsB, int32_t nc,
                              const int32_t* legsA, const int32_t* legsB,
                              const int32_t* outperm, double* C) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr, "qtn_contract_pair_dev: ctx is NULL");
  QTN_ARG((ndA == 0 || dimsA) && (ndB == 0 || dimsB) &&
              (nc == 0 || (legsA && legsB)),
          "qtn_contract_pair_dev: NULL argument");
  qtn::PairPlan pl;
  QTN_TRY(qtn::make_plan(ndA, dimsA, ndB, dimsB, nc, legsA, legsB, outperm,
                         pl));
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  return qtn::run_plan(ctx, pl, (const double2*)A, dimsA, (const double2*)B,
                       dimsB, (double2*)C);
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_contract_pair(qtn_ctx* ctx, const double* A, int32_t ndA,
                          const int64_t* dimsA, const double* B, int32_t ndB,
                          const int64_t* dimsB, int32_t nc,
                          const int32_t* legsA, const int32_t* legsB,
                          const int32_t* outperm, double* C) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr, "qtn_contract_pair: ctx is NULL");
  QTN_ARG((ndA == 0 || dimsA) && (ndB == 0 || dimsB) &&
              (nc == 0 || (legsA && legsB)),
          "qtn_contract_pair: NULL argument");
  qtn::PairPlan pl;
  QTN_TRY(qtn::make_plan(ndA, dimsA, ndB, dimsB, nc, legsA, legsB, outperm,
                         pl));
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  int64_t nA = 1, nB = 1;
  for (int i = 0; i < ndA; ++i) nA *= dimsA[i];
  for (int i = 0; i < ndB; ++i) nB *= dimsB[i];
  const int64_t nC = pl.M * pl.N;
  void *dA, *dB, *dC;
  // Slots 8..10 hold the staged operands (0..2 are the plan's scratch).
  QTN_TRY(ctx->buffer(8, std::max<int64_t>(1, nA) * sizeof(double2), &dA));
  QTN_TRY(ctx->buffer(9, std::max<int64_t>(1, nB) * sizeof(double2), &dB));
  QTN_TRY(ctx->buffer(10, std::max<int64_t>(1, nC) * sizeof(double2), &dC));
  QTN_CHECK_HIP(hipMemcpyAsync(dA, A, nA * sizeof(double2),
                               hipMemcpyHostToDevice, ctx->stream));
  QTN_CHECK_HIP(hipMemcpyAsync(dB, B, nB * sizeof(double2),
                               hipMemcpyHostToDevice, ctx->stream));
  QTN_TRY(qtn::run_plan(ctx, pl, (const double2*)dA, dimsA,
                        (const double2*)dB, dimsB, (double2*)dC));
  QTN_CHECK_HIP(hipMemcpyAsync(C, dC, nC * sizeof(double2),
                               hipMemcpyDeviceToHost, ctx->stream));
  QTN_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_permute_dev(qtn_ctx* ctx, const double* in, int32_t nd,
                        const int64_t* dims, const int32_t* perm,
                        double* out) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr, "qtn_permute_dev: ctx is NULL");
  QTN_ARG(nd >= 0 && nd <= 16, "qtn_permute_dev: rank must be 0..16");
  QTN_ARG(nd == 0 || (dims && perm), "qtn_permute_dev: NULL argument");
  std::vector<int> p0(nd);
  std::vector<char> seen(nd, 0);
  for (int i = 0; i < nd; ++i) {
    p0[i] = perm[i] - 1;
    QTN_ARG(p0[i] >= 0 && p0[i] < nd && !seen[p0[i]],
            "qtn_permute_dev: perm is not a permutation");
    seen[p0[i]] = 1;
    QTN_ARG(dims[i] >= 0, "qtn_permute_dev: negative dimension");
  }
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  if (nd == 0) {
    QTN_CHECK_HIP(hipMemcpyAsync(out, in, sizeof(double2),
                                 hipMemcpyDeviceToDevice, ctx->stream));
    return QTN_OK;
  }
  return qtn::launch_permute(ctx->stream, (const double2*)in, nd, dims,
                             p0.data(), (double2*)out);
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_permute(qtn_ctx* ctx, const double* in, int32_t nd,
                    const int64_t* dims, const int32_t* perm, double* out) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr, "qtn_permute: ctx is NULL");
  QTN_ARG(nd >= 0 && nd <= 16, "qtn_permute: rank must be 0..16");
  QTN_ARG(nd == 0 || (dims && perm), "qtn_permute: NULL argument");
  int64_t total = 1;
  for (int i = 0; i < nd; ++i) {
    QTN_ARG(dims[i] >= 0, "qtn_permute: negative dimension");
    total *= dims[i];
  }
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  void *din, *dout;
  QTN_TRY(ctx->buffer(8, std::max<int64_t>(1, total) * sizeof(double2), &din));
  QTN_TRY(ctx->buffer(9, std::max<int64_t>(1, total) * sizeof(double2),
                      &dout));
  QTN_CHECK_HIP(hipMemcpyAsync(din, in, total * sizeof(double2),
                               hipMemcpyHostToDevice, ctx->stream));
  QTN_TRY(qtn_permute_dev(ctx, (const double*)din, nd, dims, perm,
                          (double*)dout));
  QTN_CHECK_HIP(hipMemcpyAsync(out, dout, total * sizeof(double2),
                               hipMemcpyDeviceToHost, ctx->stream));
  QTN_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

}  // extern "C"
