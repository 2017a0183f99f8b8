// Batched blocked Householder QR (device side in qtn_qr.hip).
#pragma once
#include <vector>

#include "qtn_internal.hpp"

namespace qtn {

struct QrDims {
  int64_t rows;  // rows of the problem (<= ld)
  int64_t cols;  // columns (<= rows)
  int64_t active;
};

struct BatchedQR {
  int slot_base = 50;  // context scratch slots slot_base .. slot_base+7
  int batch = 0;
  int npanel = 0;
  int64_t ld = 0, stride = 0, cols_pad = 0;
  bool kept_q = false;
  QrDims* ddims = nullptr;
  double2 *V = nullptr, *VTh = nullptr, *VTn = nullptr, *Wb = nullptr;
  GemmDesc<double>* gdesc = nullptr;

  // In-place QR of every active A (rows x cols, ld lda, batch stride sA):
  // R on and above the diagonal.  keep_q keeps every panel's V and V T so
  // apply_q can form Q U afterwards.
  int32_t factor(qtn_ctx* ctx, double2* A, int64_t lda, int64_t sA,
                 int64_t cols_p, const std::vector<QrDims>& dims,
                 bool keep_q);
  // U <- Q U for every problem with ucols[b] > 0 (U rows = dims.rows,
  // ld ldu[b], batch stride sU).
  int32_t apply_q(qtn_ctx* ctx, double2* U, const std::vector<int64_t>& ldu,
                  int64_t sU, const std::vector<int64_t>& ucols,
                  const std::vector<QrDims>& dims);
};

// dst = src with its columns sorted by decreasing norm (rows_p x cols_p per
// problem, batch stride s).  perm_out receives the device permutation.
int32_t qr_sort_columns(qtn_ctx* ctx, const double2* src, double2* dst,
                        int64_t rows_p, int64_t cols_p, int64_t s, int batch,
                        const QrDims* d_dims, int slot_norm, int slot_perm,
                        int** perm_out);
// Y = R^H (lower triangular) from the R factor left in A by factor().
int32_t qr_r_to_adjoint(qtn_ctx* ctx, const double2* A, int64_t lda,
                        int64_t sA, double2* Y, int64_t ldy, int64_t sY,
                        int64_t ny_rows, int64_t ny_cols, int batch,
                        const QrDims* d_dims);

}  // namespace qtn
