// Batched block one-sided Jacobi SVD engine (device side lives in
// qtn_svd.hip).  The caller owns X (and optionally W); the engine owns the
// Gram partials, rotation blocks and the per-problem result arrays, all in
// context scratch slots.
#pragma once
#include <vector>

#include "qtn_internal.hpp"

namespace qtn {

// Context scratch slots used by the engine (callers use other slots).
constexpr int kSlotGram = 20;
constexpr int kSlotQ = 21;
constexpr int kSlotSmall = 22;
constexpr int kSlotSig = 23;
constexpr int kSlotPerm = 24;

struct JacobiEngine {
  int batch = 0;
  int64_t rows_p = 0;  // padded rows of X (multiple of 64)
  int64_t cols_p = 0;  // padded columns of X (multiple of 64)
  int W = 64;          // block-pair width (columns of one Gram block)
  int p = 0;           // block pairs per round (= cols_p / W)
  bool single = false; // FP32 sweeps (complex float X, Gram and rotations)
  int slot_base = kSlotGram;  // first of the 5 context slots it uses
  bool fused = false;  // W = 32 rounds as one jac_round launch
  // FP32 fused rounds whose X (and W) tiles fit the LDS-resident kernel
  // (at most 512 rows) run jac_round_res; the other fused rounds jac_round
  bool res_round(int nchunk_x, int nchunk_w) const {
    return single && nchunk_x * 64 <= 512 && nchunk_w * 64 <= 512;
  }
  // the inner sweeps of outer rounds 1.. rotate only the cross pairs of
  // their two column blocks (round 0 of every outer sweep rotates all)
  bool cross_rounds = false;
  int64_t ldw = 0;     // leading dimension of W (0: cols_p)
  const int64_t* chimax_per = nullptr;  // host, per-problem cap (optional)
  int64_t* chimax_dev = nullptr;
  // device arrays (valid after reserve); Gram partials and rotation blocks
  // are complex double, or complex float when `single`
  void* Gp = nullptr;
  void* Q = nullptr;
  double* normF2 = nullptr;
  double* zthr = nullptr;
  double* disc = nullptr;   // discarded weight per problem
  double* total = nullptr;  // sum of all sigma^2 per problem
  double* keptw = nullptr;  // sum of the kept sigma^2 per problem
  double* offmax = nullptr; // largest scaled off-diagonal of the sweep
  int64_t* kept = nullptr;
  int64_t* nsv = nullptr;
  int* flags = nullptr;
  int* counters = nullptr;
  int* active = nullptr;
  int* status = nullptr;
  int* want_w = nullptr;
  int* mask = nullptr;  // problems to solve (host mask of run*)
  int* scale_e = nullptr;  // X is swept as X * 2^-scale_e (jac_prep_final)
  double* prep_scale = nullptr;  // per-block partial norms (jac_prep_*)
  double* prep_ssq = nullptr;
  int* prep_bad = nullptr;
  int* misc = nullptr;
  double* sig = nullptr;  // [batch][cols_p], sorted descending after run
  qtn_ctx* ctx_ = nullptr;
  std::vector<hipEvent_t> prof_events;  // 6 per round while profiling
  int32_t account_sweep(int n_active, int qpairs, int qpairs_w, int nchunk,
                        int nchunk_w);
  int* perm = nullptr;    // [batch][cols_p], column of X for each sigma

  // n_active: problems that will actually be solved (the rest are masked;
  // < 0: all of them); it selects the tuned tiling's batch class.
  int32_t reserve(qtn_ctx* ctx, int batch, int64_t rows_p, int64_t cols_p,
                  bool single = false, int n_active = -1);
  // Orthogonalise the columns of every X (rows_p x cols_p, ld rows_p,
  // batch stride sX); if W is given (cols_p x cols_p, stride sW) it
  // accumulates the rotations starting from the identity, for the problems
  // with want_w_host[b] != 0 (all of them when want_w_host is NULL).
  // Afterwards sig / perm hold the sorted column norms and kept / disc /
  // total / status the truncation decision for nsv_host[b] meaningful
  // values.
  template <typename R, int WW>
  int32_t launch_round(hipStream_t s, void* X, int64_t ldx, int64_t sX,
                       void* Wm, int64_t sW, int r, double tol, int n_inner,
                       int nchunk_x, int nchunk_w, bool per_problem_w);
  template <typename R>
  int32_t run_t(qtn_ctx* ctx, typename CType<R>::type* X, int64_t ldx,
                int64_t sX, typename CType<R>::type* Wm, int64_t sW,
                const int* want_w_host, const int64_t* nsv_host,
                double cutoff, int relative, int64_t chimax, int max_sweeps,
                int* sweeps_done, const int* mask_host);
  // X: rows_p rows are processed, leading dimension ldx >= rows_p.
  // mask_host (may be NULL = all): problems with mask 0 are left alone and
  // their results are meaningless.
  int32_t run(qtn_ctx* ctx, double2* X, int64_t ldx, int64_t sX, double2* Wm,
              int64_t sW,
              const int* want_w_host, const int64_t* nsv_host, double cutoff,
              int relative, int64_t chimax, int max_sweeps, int* sweeps_done,
              const int* mask_host = nullptr);
  // The same sweeps in complex FP32 (engine reserved with single = true).
  // Columns that end below the FP32 noise floor (8 sqrt(rows) eps32 |X|_F)
  // are never rotated and count as exact zeros in sig and the truncation.
  int32_t run32(qtn_ctx* ctx, float2* X, int64_t ldx, int64_t sX, float2* Wm,
                int64_t sW, const int* want_w_host, const int64_t* nsv_host,
                double cutoff, int relative, int64_t chimax, int max_sweeps,
                int* sweeps_done, const int* mask_host = nullptr);
};

int32_t launch_identity(hipStream_t s, double2* W, int64_t n, int64_t sW,
                        int batch);
template <typename R>
int32_t launch_identity_t(hipStream_t s, typename CType<R>::type* W,
                          int64_t n, int64_t ld, int64_t sW, int batch);

inline int64_t round_up64(int64_t x) { return ((x + 63) / 64) * 64; }

// Batch class of the tuned tilings: floor(log2(active problems)), 0..6
// (1, 2-3, 4-7, ..., 64 and more).
inline int batch_class(int n) {
  int c = 0;
  while (c < 6 && (2 << c) <= n) ++c;
  return c;
}

// SVD of a device matrix (defined in qtn_svd_api.hip; contract documented
// there).  Uses context slots 12 and 13, and 118-139 for the
// QR-preconditioned sweeps of splits with 512 or more columns (opt-in with
// QTN_SVD_PRECOND=1, and only when allow_pre).
int32_t svd_device(qtn_ctx* ctx, const double2* dA, int64_t m, int64_t n,
                   int64_t lda, bool truncate, double cutoff, int relative,
                   int64_t chimax, int absorb, double2* dU, double* dS,
                   double2* dVh, int64_t* k_out, double* disc_out,
                   bool allow_pre = true);
// Left singular vectors and values of A (adj == 0) or A^H (adj != 0) only
// (the caller forms the other factor by a GEMM); contract in
// qtn_svd_api.hip.  Uses slots 12, 15 and 118-139.
int32_t svd_device_left(qtn_ctx* ctx, const double2* dA, int64_t m,
                        int64_t n, int64_t lda, int adj, double cutoff,
                        int relative, int64_t chimax, double2* dU, double* dS,
                        int64_t* k_out, double* disc_out);
// The chosen form (FP32 sweeps where the cap decides the truncation, FP64
// re-orthonormalisation and recomputation); contract in qtn_svd_api.hip.
int32_t svd_device_chosen(qtn_ctx* ctx, const double2* dA, int64_t m,
                          int64_t n, int64_t lda, double cutoff, int relative,
                          int64_t chimax, int absorb, double2* dU, double* dS,
                          double2* dVh, int64_t* k_out, double* disc_out,
                          bool* fallback);

}  // namespace qtn
