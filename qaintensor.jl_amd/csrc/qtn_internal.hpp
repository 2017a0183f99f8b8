// Internal declarations shared by the libqtn translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "qtn.h"

namespace qtn {

void set_error(const std::string& msg);
void clear_error();
// Status for an exception caught at the C ABI (or in a worker thread):
// QTN_ERR_OOM for a failed host allocation, QTN_ERR_HIP otherwise, with
// the exception's text as the error message.  Call inside a catch block.
int32_t abi_exception() noexcept;

#define QTN_CHECK_HIP(call)                                                  \
  do {                                                                       \
    hipError_t qtn_e_ = (call);                                              \
    if (qtn_e_ != hipSuccess) {                                              \
      ::qtn::set_error(std::string(#call) + " failed: " +                    \
                       hipGetErrorString(qtn_e_));                           \
      return QTN_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

#define QTN_TRY(expr)                                                        \
  do {                                                                       \
    int32_t qtn_s_ = (expr);                                                 \
    if (qtn_s_ != QTN_OK) return qtn_s_;                                     \
  } while (0)

#define QTN_ARG(cond, msg)                                                   \
  do {                                                                       \
    if (!(cond)) {                                                           \
      ::qtn::set_error(msg);                                                 \
      return QTN_ERR_ARG;                                                    \
    }                                                                        \
  } while (0)

// A growable device buffer owned by the context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// Complex GEMM launch (device pointers, element counts; complex elements).
// op: 0 = N, 1 = T, 2 = C.  Returns QTN_OK or an error status.
int32_t launch_zgemm(hipStream_t s, int opA, int opB, int64_t m, int64_t n,
                     int64_t k, double2 alpha, const double2* A, int64_t lda,
                     int64_t sA, const double2* B, int64_t ldb, int64_t sB,
                     double2 beta, double2* C, int64_t ldc, int64_t sC,
                     int64_t batch);
int32_t launch_cgemm(hipStream_t s, int opA, int opB, int64_t m, int64_t n,
                     int64_t k, float2 alpha, const float2* A, int64_t lda,
                     int64_t sA, const float2* B, int64_t ldb, int64_t sB,
                     float2 beta, float2* C, int64_t ldc, int64_t sC,
                     int64_t batch);
// Per-problem descriptor of a grouped (uneven batched) complex GEMM.
template <typename R>
struct CType;
template <>
struct CType<double> {
  typedef double2 type;
};
template <>
struct CType<float> {
  typedef float2 type;
};
template <typename R>
struct GemmDesc {
  int64_t M, N, K;
  typename CType<R>::type alpha, beta;
  const typename CType<R>::type* A;
  int64_t lda;
  const typename CType<R>::type* B;
  int64_t ldb;
  typename CType<R>::type* C;
  int64_t ldc;
};
// d_desc: device array of `count` descriptors; maxM / maxN bound every
// problem's M / N (they size the grid).
int32_t launch_zgemm_grouped(hipStream_t s, int opA, int opB,
                             const GemmDesc<double>* d_desc, int count,
                             int64_t maxM, int64_t maxN);
int32_t launch_cgemm_grouped(hipStream_t s, int opA, int opB,
                             const GemmDesc<float>* d_desc, int count,
                             int64_t maxM, int64_t maxN);
int32_t launch_permute(hipStream_t s, const double2* in, int nd,
                       const int64_t* dims, const int* perm0, double2* out);

// BF16x3 split-precision complex GEMM (qtn_bf16x3.hip): C = op(A) B, opA
// 0 (N) or 2 (C), C overwritten; scratch in context slots 190-191.
int32_t launch_zgemm_bf16x3(qtn_ctx* ctx, int opA, int64_t m, int64_t n,
                            int64_t k, const double2* A, int64_t lda,
                            const double2* B, int64_t ldb, double2* C,
                            int64_t ldc);

// Looks up this GPU's copy of the no-allocation scratch (qtn_contract.hip);
// called once per device with it current (qtn_ctx_create).
int32_t noalloc_scratch_init(int device);

// Rank-predictor MLP (qtn_mlp.hip): prediction on device features, Y
// points into the MLP's own scratch.
int32_t mlp_predict_device(qtn_mlp* m, int64_t nb, const float* X,
                           float** Y);
int mlp_nin(const qtn_mlp* m);
qtn_ctx* mlp_ctx(const qtn_mlp* m);

}  // namespace qtn

struct qtn_ctx {
  int device = 0;
  uint64_t sketch_seed = 0x5eed;  // range-finder sketches (learned form)
  hipStream_t stream = nullptr;
  // Second stream (default priority) and two events for work that overlaps
  // the main stream inside one call, e.g. the trailing updates of a QR
  // behind its next panel; created on first use (side_stream()).
  hipStream_t side = nullptr;
  hipEvent_t side_ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int32_t side_stream(hipStream_t* out);
  // Instantiated graphs of one Jacobi sweep, keyed by the bytes of every
  // value its launches take (sweep_graph captures on a miss, evicting the
  // oldest beyond kMaxSweepGraphs; callers have no graph in flight then).
  struct CachedGraph {
    std::vector<char> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  static constexpr size_t kMaxSweepGraphs = 32;
  std::vector<CachedGraph> sweep_graphs;
  int32_t sweep_graph(const void* key, size_t n, hipStream_t s,
                      const std::function<int32_t()>& enqueue,
                      hipGraphExec_t* out);
  // Jacobi block-pair tiling per (padded column count, FP32 sweeps, batch
  // class of the active problems): set by qtn_tune_jacobi /
  // qtn_set_jacobi_tiling (QTN_TILING_* codes).
  std::map<std::tuple<int64_t, int, int>, int> jac_tiling;
  // Named scratch buffers grown on demand (never shrunk until destroy).
  std::vector<qtn::DevBuf> bufs;
  // work counters (see qtn_ctx_counters)
  int64_t n_svd = 0, n_sweeps = 0, max_sweeps = 0, n_batches = 0;
  // chosen-form single splits (svd_device_chosen): Rayleigh-Ritz splits
  // accepted, and splits redone by the FP64 engine
  int64_t n_rr_accepted = 0, n_rr_redone = 0;
  // per-kernel event timing (qtn_ctx_profile)
  // ids: 0 jac_inner, 1 jac_gram, 2 jac_update, 3 jac_round (FP64 sweeps),
  // 4 jac_inner, 5 jac_gram, 6 jac_update, 7 jac_round (FP32 sweeps of the
  // chosen form)
  static constexpr int kProfKernels = 8;
  bool prof = false;
  int64_t prof_count[kProfKernels] = {};
  double prof_ms[kProfKernels] = {};
  double prof_bytes[kProfKernels] = {};
  double prof_flops[kProfKernels] = {};
  std::vector<hipEvent_t> ev_pool;  // recycled events
  // Small pinned host staging buffer for scalar read-backs.
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  // Returns a device buffer of at least `bytes` for slot `slot`.
  int32_t buffer(int slot, size_t bytes, void** out);
  int32_t pinned_buffer(size_t bytes, void** out);
  // Worker contexts (same device, own stream and scratch) that run
  // independent batches concurrently with this one; created on first use,
  // destroyed with it.  parent is set on workers.
  qtn_ctx* parent = nullptr;
  std::vector<qtn_ctx*> workers;
  int32_t worker(int i, qtn_ctx** out);
};

// Device-resident MPS (qtn_mps.hip): right-canonical site tensors with the
// bond spectra beside them, in one arena.
struct qtn_mps {
  qtn_ctx* ctx = nullptr;
  int n = 0;
  int64_t cap = 0;
  std::vector<int64_t> dims;     // chi_0..chi_n
  std::vector<int64_t> maxb;     // capacity of every bond
  std::vector<double2*> site;    // device site tensors (0-based site index)
  std::vector<double*> lam;      // device bond spectra, lam[j] for bond j
  void* arena = nullptr;         // one allocation for all of the above
};
