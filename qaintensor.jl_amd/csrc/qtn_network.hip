// Whole-network contraction with every intermediate resident on the device:
// the work of TensorOperations.ncon(tensors, labels; order) behind
// contract(net) (src/contract.jl:242-264; ncon at :257 and :263).
//
// Labels follow ncon: a positive label names a contracted leg pair (it
// appears twice in the network, in one tensor for a trace), a negative one
// an open leg; the result's legs are ordered -1, -2, ... .  The contraction
// sequence is the caller's (the reference's planner, or increasing labels):
// the first label not yet contracted names the next pair of tensors, which
// are contracted over every label they share.  Tensors left unconnected at
// the end are joined by outer products in their order.
//
// Each pairwise step is one launch of a complex FP64 GEMM on MFMA whose
// operand staging gathers through the tensors' own strides (the permutes
// of a permute + GEMM scheme folded into the global -> LDS loads): the
// free legs of A index the rows, the shared legs the inner dimension, the
// free legs of B the columns, and the product is written contiguous with
// legs (free A, free B).  Intermediates live in one device arena whose
// layout (first fit, freed operands reused) is planned on the host before
// anything runs; the inputs arrive in one staging copy, the result leaves
// in one.
#include <algorithm>
#include <cstring>
#include <functional>
#include <map>
#include <numeric>
#include <vector>

#include "qtn_internal.hpp"

namespace qtn {
namespace {

constexpr int kMaxLegs = 24;
constexpr int kSlotArena = 220;

// A group of legs seen as one flat index (first leg fastest) and the
// element offset each of its values adds.
struct LegSet {
  int32_t n;
  int32_t dim[kMaxLegs];
  int64_t stride[kMaxLegs];
};

__device__ __forceinline__ int64_t leg_offset(const LegSet& s, int64_t idx) {
  int64_t off = 0;
  if (idx <= 0x7fffffff) {
    // 32-bit divisions where the index allows (64-bit ones are long
    // instruction sequences)
    uint32_t u = (uint32_t)idx;
    for (int i = 0; i < s.n; ++i) {
      const uint32_t d = (uint32_t)s.dim[i];
      const uint32_t q = u / d;
      off += (int64_t)(u - q * d) * s.stride[i];
      u = q;
    }
    return off;
  }
  for (int i = 0; i < s.n; ++i) {
    const int64_t d = s.dim[i];
    const int64_t q = idx / d;
    off += (idx - q * d) * s.stride[i];
    idx = q;
  }
  return off;
}

typedef double d4_t __attribute__((ext_vector_type(4)));

// C (M x N, contiguous) = sum_k A[am(m) + ak(k)] B[bk(k) + bn(n)].
// 256 threads, tile (32 TM) x (32 TN) x 16, four waves of (16 TM) x (16 TN)
// on v_mfma_f64_16x16x4f64 (four real products per complex one).  The
// tile's row and column offsets and each k-slice's offsets are computed
// once into LDS; a_kfast / b_nfast pick the staging order that reads the
// operand's unit-stride leg along consecutive threads.
template <int TM, int TN>
__global__ __launch_bounds__(256) void net_gemm(
    const double2* __restrict__ A, const double2* __restrict__ B,
    double2* __restrict__ C, int64_t M, int64_t N, int64_t K, LegSet am,
    LegSet ak, LegSet bk, LegSet bn, int a_kfast, int b_nfast) {
  constexpr int BM = 32 * TM, BN = 32 * TN, BK = 16;
  __shared__ double As_re[BK][BM + 1], As_im[BK][BM + 1];
  __shared__ double Bs_re[BK][BN + 1], Bs_im[BK][BN + 1];
  __shared__ int64_t offAm[BM], offBn[BN], offAk[BK], offBk[BK];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
  for (int i = tid; i < BM; i += 256)
    offAm[i] = m0 + i < M ? leg_offset(am, m0 + i) : 0;
  for (int j = tid; j < BN; j += 256)
    offBn[j] = n0 + j < N ? leg_offset(bn, n0 + j) : 0;
  const int wm = (wave & 1) * TM * 16, wn = (wave >> 1) * TN * 16;
  d4_t acc_re[TM][TN], acc_im[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        acc_re[i][j][e] = 0.0;
        acc_im[i][j][e] = 0.0;
      }
  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();  // the previous slice's MFMA reads are done
    if (tid < BK)
      offAk[tid] = k0 + tid < K ? leg_offset(ak, k0 + tid) : 0;
    else if (tid < 2 * BK)
      offBk[tid - BK] = k0 + tid - BK < K ? leg_offset(bk, k0 + tid - BK) : 0;
    __syncthreads();
    if (!a_kfast) {
      const int i = tid % BM;
      for (int kk = tid / BM; kk < BK; kk += 256 / BM) {
        double2 v = make_double2(0.0, 0.0);
        if (m0 + i < M && k0 + kk < K) v = A[offAm[i] + offAk[kk]];
        As_re[kk][i] = v.x;
        As_im[kk][i] = v.y;
      }
    } else {
      const int kk = tid % BK;
      for (int i = tid / BK; i < BM; i += 256 / BK) {
        double2 v = make_double2(0.0, 0.0);
        if (m0 + i < M && k0 + kk < K) v = A[offAm[i] + offAk[kk]];
        As_re[kk][i] = v.x;
        As_im[kk][i] = v.y;
      }
    }
    if (b_nfast) {
      const int j = tid % BN;
      for (int kk = tid / BN; kk < BK; kk += 256 / BN) {
        double2 v = make_double2(0.0, 0.0);
        if (n0 + j < N && k0 + kk < K) v = B[offBk[kk] + offBn[j]];
        Bs_re[kk][j] = v.x;
        Bs_im[kk][j] = v.y;
      }
    } else {
      const int kk = tid % BK;
      for (int j = tid / BK; j < BN; j += 256 / BK) {
        double2 v = make_double2(0.0, 0.0);
        if (n0 + j < N && k0 + kk < K) v = B[offBk[kk] + offBn[j]];
        Bs_re[kk][j] = v.x;
        Bs_im[kk][j] = v.y;
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4);
      double a_re[TM], a_im[TM], b_re[TN], b_im[TN];
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
          acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              a_re[i], b_re[j], acc_re[i][j], 0, 0, 0);
          acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              -a_im[i], b_im[j], acc_re[i][j], 0, 0, 0);
          acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              a_re[i], b_im[j], acc_im[i][j], 0, 0, 0);
          acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              a_im[i], b_re[j], acc_im[i][j], 0, 0, 0);
        }
    }
  }
  // result element e of lane l: row 4 e + l / 16, column l % 16
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int64_t col = n0 + wn + j * 16 + (lane & 15);
      if (col >= N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = m0 + wm + i * 16 + 4 * e + (lane >> 4);
        if (row < M)
          C[row + col * M] = make_double2(acc_re[i][j][e], acc_im[i][j][e]);
      }
    }
}

// out[o] = in[src(o)] over `total` elements (the final leg order).
__global__ __launch_bounds__(256) void net_gather(
    const double2* __restrict__ in, double2* __restrict__ out, int64_t total,
    LegSet src) {
  for (int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x; o < total;
       o += (int64_t)gridDim.x * 256)
    out[o] = in[leg_offset(src, o)];
}

// d x d identity (a trace is a contraction with it).
__global__ void net_identity(double2* __restrict__ out, int64_t d) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < d * d;
       i += (int64_t)gridDim.x * 256)
    out[i] = make_double2((i % d) == (i / d) ? 1.0 : 0.0, 0.0);
}

struct Item {
  std::vector<int64_t> labels, dims;
  int64_t off = 0, size = 1;  // in elements
};

std::vector<int64_t> col_strides(const std::vector<int64_t>& dims) {
  std::vector<int64_t> s(dims.size());
  int64_t acc = 1;
  for (size_t i = 0; i < dims.size(); ++i) {
    s[i] = acc;
    acc *= dims[i];
  }
  return s;
}

// Legs (dims, strides) -> LegSet, merging neighbours that are contiguous in
// memory.  False when more than kMaxLegs groups remain.
bool make_legset(const std::vector<int64_t>& dims,
                 const std::vector<int64_t>& strides, LegSet* out) {
  std::vector<int64_t> d, st;
  for (size_t i = 0; i < dims.size(); ++i) {
    if (dims[i] == 1) continue;
    if (!d.empty() && st.back() * d.back() == strides[i]) {
      d.back() *= dims[i];
    } else {
      d.push_back(dims[i]);
      st.push_back(strides[i]);
    }
  }
  if ((int)d.size() > kMaxLegs) return false;
  std::memset(out, 0, sizeof(LegSet));
  out->n = (int32_t)d.size();
  for (size_t i = 0; i < d.size(); ++i) {
    if (d[i] > INT32_MAX) return false;
    out->dim[i] = (int32_t)d[i];
    out->stride[i] = st[i];
  }
  return true;
}

// The inner index space is shared by A and B: neighbours merge only where
// they are contiguous in both.
bool make_kpair(const std::vector<int64_t>& dims,
                const std::vector<int64_t>& sa, const std::vector<int64_t>& sb,
                LegSet* oa, LegSet* ob) {
  std::vector<int64_t> d, ta, tb;
  for (size_t i = 0; i < dims.size(); ++i) {
    if (dims[i] == 1) continue;
    if (!d.empty() && ta.back() * d.back() == sa[i] &&
        tb.back() * d.back() == sb[i]) {
      d.back() *= dims[i];
    } else {
      d.push_back(dims[i]);
      ta.push_back(sa[i]);
      tb.push_back(sb[i]);
    }
  }
  if ((int)d.size() > kMaxLegs) return false;
  std::memset(oa, 0, sizeof(LegSet));
  std::memset(ob, 0, sizeof(LegSet));
  oa->n = ob->n = (int32_t)d.size();
  for (size_t i = 0; i < d.size(); ++i) {
    if (d[i] > INT32_MAX) return false;
    oa->dim[i] = ob->dim[i] = (int32_t)d[i];
    oa->stride[i] = ta[i];
    ob->stride[i] = tb[i];
  }
  return true;
}

// First-fit placement in an arena of elements (offsets 16-aligned).
struct Arena {
  std::map<int64_t, int64_t> used;  // offset -> size
  int64_t peak = 0;
  int64_t alloc(int64_t size) {
    const int64_t sz = std::max<int64_t>(16, (size + 15) / 16 * 16);
    int64_t at = 0;
    for (const auto& u : used) {
      if (u.first - at >= sz) break;
      at = std::max(at, u.first + (u.second + 15) / 16 * 16);
    }
    used[at] = sz;
    peak = std::max(peak, at + sz);
    return at;
  }
  void release(int64_t off) { used.erase(off); }
};

// One step of the planned execution.
struct Step {
  int kind;          // 0 pairwise GEMM, 1 identity fill
  int a = -1, b = -1, out = -1;
  int64_t M = 1, N = 1, K = 1;
  LegSet am, ak, bk, bn;
  int a_kfast = 0, b_nfast = 0;
};

int64_t prod(const std::vector<int64_t>& v) {
  int64_t p = 1;
  for (int64_t x : v) p *= x;
  return p;
}

}  // namespace
}  // namespace qtn

using namespace qtn;

extern "C" {

int32_t qtn_contract_network(qtn_ctx* ctx, int32_t ntensors,
                             const int32_t* ndims, const int64_t* dims,
                             const int32_t* labels,
                             const double* const* tensors, int32_t norder,
                             const int32_t* order, double* out,
                             int64_t out_len) try {
  clear_error();
  QTN_ARG(ctx && ntensors >= 1 && ndims && tensors && out,
          "qtn_contract_network: NULL argument or no tensors");
  QTN_ARG(norder >= 0 && (norder == 0 || order),
          "qtn_contract_network: bad contraction order");
  // ---- read the network
  std::vector<Item> items(ntensors);
  std::map<int64_t, int> count;
  std::map<int64_t, int64_t> ldim;
  int64_t pos = 0;
  for (int t = 0; t < ntensors; ++t) {
    QTN_ARG(ndims[t] >= 0, "qtn_contract_network: negative rank");
    QTN_ARG(tensors[t], "qtn_contract_network: NULL tensor");
    for (int l = 0; l < ndims[t]; ++l, ++pos) {
      const int64_t lab = labels[pos], d = dims[pos];
      QTN_ARG(lab != 0, "qtn_contract_network: label 0 is not allowed");
      QTN_ARG(d >= 1, "qtn_contract_network: leg dimensions must be >= 1");
      items[t].labels.push_back(lab);
      items[t].dims.push_back(d);
      count[lab] += 1;
      if (ldim.count(lab))
        QTN_ARG(ldim[lab] == d,
                "Dimensions of contraction legs do not match");
      ldim[lab] = d;
    }
    items[t].size = prod(items[t].dims);
  }
  std::vector<int64_t> open;
  for (const auto& kv : count) {
    if (kv.first > 0)
      QTN_ARG(kv.second == 2, "qtn_contract_network: a positive label must "
                              "appear exactly twice");
    else {
      QTN_ARG(kv.second == 1, "qtn_contract_network: a negative label must "
                              "appear once");
      open.push_back(kv.first);
    }
  }
  {
    std::vector<int64_t> o(order, order + norder), p;
    for (const auto& kv : count)
      if (kv.first > 0) p.push_back(kv.first);
    std::sort(o.begin(), o.end());
    QTN_ARG(o == p, "qtn_contract_network: the order must list every "
                    "positive label once");
  }
  // ncon's output order: -1, -2, ...
  std::sort(open.begin(), open.end(), std::greater<int64_t>());
  int64_t out_size = 1;
  for (int64_t lab : open) out_size *= ldim[lab];
  QTN_ARG(out_len == out_size,
          "qtn_contract_network: out_len differs from the result's size");

  // ---- plan: steps and arena placement (host only)
  Arena arena;
  std::vector<int> alive;
  int64_t in_total = 0;
  for (int t = 0; t < ntensors; ++t) {
    items[t].off = arena.alloc(items[t].size);
    in_total = std::max(in_total, items[t].off + items[t].size);
    alive.push_back(t);
  }
  std::vector<Step> steps;
  int64_t next_label = 1;
  for (const auto& kv : count) next_label = std::max(next_label, kv.first + 1);
  auto pair_step = [&](int ia, int ib) -> int32_t {
    Item &A = items[ia], &B = items[ib];
    std::vector<int64_t> shared;
    for (int64_t x : A.labels)
      if (x > 0 &&
          std::find(B.labels.begin(), B.labels.end(), x) != B.labels.end())
        shared.push_back(x);
    const std::vector<int64_t> sA = col_strides(A.dims),
                               sB = col_strides(B.dims);
    // inner legs follow the larger operand's leg order
    if (B.size > A.size) {
      std::vector<int64_t> sh;
      for (int64_t x : B.labels)
        if (std::find(shared.begin(), shared.end(), x) != shared.end())
          sh.push_back(x);
      shared = sh;
    }
    Item C;
    std::vector<int64_t> dm, sm, dn, sn, dk, ska, skb;
    for (size_t i = 0; i < A.labels.size(); ++i)
      if (std::find(shared.begin(), shared.end(), A.labels[i]) ==
          shared.end()) {
        C.labels.push_back(A.labels[i]);
        C.dims.push_back(A.dims[i]);
        dm.push_back(A.dims[i]);
        sm.push_back(sA[i]);
      }
    for (size_t i = 0; i < B.labels.size(); ++i)
      if (std::find(shared.begin(), shared.end(), B.labels[i]) ==
          shared.end()) {
        C.labels.push_back(B.labels[i]);
        C.dims.push_back(B.dims[i]);
        dn.push_back(B.dims[i]);
        sn.push_back(sB[i]);
      }
    for (int64_t x : shared) {
      const size_t ia_ = std::find(A.labels.begin(), A.labels.end(), x) -
                         A.labels.begin();
      const size_t ib_ = std::find(B.labels.begin(), B.labels.end(), x) -
                         B.labels.begin();
      dk.push_back(A.dims[ia_]);
      ska.push_back(sA[ia_]);
      skb.push_back(sB[ib_]);
    }
    Step st;
    st.kind = 0;
    st.a = ia;
    st.b = ib;
    st.M = prod(dm);
    st.N = prod(dn);
    st.K = prod(dk);
    if (!make_legset(dm, sm, &st.am) || !make_legset(dn, sn, &st.bn) ||
        !make_kpair(dk, ska, skb, &st.ak, &st.bk)) {
      set_error("qtn_contract_network: a step has more than 24 leg groups");
      return QTN_ERR_ARG;
    }
    // stage along the operand's unit-stride leg where it is in the group
    st.a_kfast = (st.ak.n > 0 && st.ak.stride[0] == 1) ? 1 : 0;
    st.b_nfast = (st.bn.n > 0 && st.bn.stride[0] == 1) ? 1 : 0;
    C.size = st.M * st.N;
    C.off = arena.alloc(C.size);
    items.push_back(C);
    st.out = (int)items.size() - 1;
    arena.release(items[ia].off);
    arena.release(items[ib].off);
    alive.erase(std::find(alive.begin(), alive.end(), ia));
    alive.erase(std::find(alive.begin(), alive.end(), ib));
    alive.push_back(st.out);
    steps.push_back(st);
    return QTN_OK;
  };
  std::map<int64_t, bool> done;
  for (int q = 0; q < norder; ++q) {
    const int64_t lab = order[q];
    if (done[lab]) continue;
    std::vector<int> holders;
    for (int t : alive)
      if (std::find(items[t].labels.begin(), items[t].labels.end(), lab) !=
          items[t].labels.end())
        holders.push_back(t);
    if (holders.size() == 1) {
      // a trace: the second leg takes a fresh label and the tensor meets
      // the identity on both
      Item& A = items[holders[0]];
      auto it = std::find(A.labels.begin(), A.labels.end(), lab);
      it = std::find(it + 1, A.labels.end(), lab);
      const int64_t fresh = next_label++;
      *it = fresh;
      Item I;
      I.labels = {lab, fresh};
      I.dims = {ldim[lab], ldim[lab]};
      I.size = ldim[lab] * ldim[lab];
      I.off = arena.alloc(I.size);
      items.push_back(I);
      Step f;
      f.kind = 1;
      f.out = (int)items.size() - 1;
      steps.push_back(f);
      alive.push_back(f.out);
      holders.push_back(f.out);
      done[fresh] = true;
    }
    QTN_ARG(holders.size() == 2, "qtn_contract_network: inconsistent labels");
    for (int64_t x : items[holders[0]].labels)
      if (x > 0 && std::find(items[holders[1]].labels.begin(),
                             items[holders[1]].labels.end(),
                             x) != items[holders[1]].labels.end())
        done[x] = true;
    QTN_TRY(pair_step(holders[0], holders[1]));
  }
  while (alive.size() > 1) QTN_TRY(pair_step(alive[0], alive[1]));
  const Item& R = items[alive[0]];
  // final gather into ncon's leg order
  LegSet fin;
  {
    const std::vector<int64_t> sR = col_strides(R.dims);
    std::vector<int64_t> d, s;
    for (int64_t lab : open) {
      const size_t i =
          std::find(R.labels.begin(), R.labels.end(), lab) - R.labels.begin();
      d.push_back(R.dims[i]);
      s.push_back(sR[i]);
    }
    QTN_ARG(make_legset(d, s, &fin),
            "qtn_contract_network: the result has more than 24 leg groups");
  }
  const int64_t fin_off = arena.alloc(out_size);

  // ---- run
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  void* base;
  QTN_TRY(ctx->buffer(kSlotArena, (size_t)arena.peak * sizeof(double2),
                      &base));
  double2* dev = (double2*)base;
  // inputs through one pinned staging copy
  void* pin = nullptr;
  QTN_TRY(ctx->pinned_buffer((size_t)in_total * sizeof(double2), &pin));
  for (int t = 0; t < ntensors; ++t)
    std::memcpy((double2*)pin + items[t].off, tensors[t],
                (size_t)items[t].size * sizeof(double2));
  QTN_CHECK_HIP(hipMemcpyAsync(dev, pin, (size_t)in_total * sizeof(double2),
                               hipMemcpyHostToDevice, s));
  for (const Step& st : steps) {
    if (st.kind == 1) {
      const Item& I = items[st.out];
      hipLaunchKernelGGL(net_identity, dim3(64), dim3(256), 0, s,
                         dev + I.off, I.dims[0]);
      continue;
    }
    const double2* A = dev + items[st.a].off;
    const double2* B = dev + items[st.b].off;
    double2* C = dev + items[st.out].off;
    const int tm = st.M > 32 ? 2 : 1, tn = st.N > 32 ? 2 : 1;
    const int64_t gx = (st.M + 32 * tm - 1) / (32 * tm);
    const int64_t gy = (st.N + 32 * tn - 1) / (32 * tn);
    QTN_ARG(gx <= INT32_MAX && gy <= 65535,
            "qtn_contract_network: a step is too large for the launch grid");
    const dim3 grid((unsigned)gx, (unsigned)gy);
#define QTN_NET_GEMM(TM_, TN_)                                             \
  hipLaunchKernelGGL((net_gemm<TM_, TN_>), grid, dim3(256), 0, s, A, B, C, \
                     st.M, st.N, st.K, st.am, st.ak, st.bk, st.bn,         \
                     st.a_kfast, st.b_nfast)
    if (tm == 2 && tn == 2)
      QTN_NET_GEMM(2, 2);
    else if (tm == 2)
      QTN_NET_GEMM(2, 1);
    else if (tn == 2)
      QTN_NET_GEMM(1, 2);
    else
      QTN_NET_GEMM(1, 1);
#undef QTN_NET_GEMM
  }
  QTN_CHECK_HIP(hipGetLastError());
  const unsigned gg = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>((out_size + 255) / 256, 4096));
  hipLaunchKernelGGL(net_gather, dim3(gg), dim3(256), 0, s,
                     (const double2*)(dev + R.off), dev + fin_off, out_size,
                     fin);
  QTN_CHECK_HIP(hipGetLastError());
  QTN_CHECK_HIP(hipMemcpyAsync(out, dev + fin_off,
                               (size_t)out_size * sizeof(double2),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

}  // extern "C"
