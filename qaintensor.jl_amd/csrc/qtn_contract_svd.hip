// contract_svd (src/svd.jl:7-38): contract two tensors over one leg after
// truncating each operand's SVD at the absolute tail norm er.  Everything
// runs on the device: two permutes, two Jacobi SVDs, three complex GEMMs.
#include <algorithm>
#include <vector>

#include "qtn_internal.hpp"
#include "qtn_svd.hpp"

namespace qtn {
namespace {

// Slots: 8/9 staged operands, 3/4 permuted matrices, 5/6 factors,
// 7/0 truncated reconstructions, 10 result.
int32_t trunc_reconstruct(qtn_ctx* ctx, const double2* M, int64_t rows,
                          int64_t cols, double er, double2* out) {
  const int64_t kk = std::min(rows, cols);
  void *dU, *dV;
  QTN_TRY(ctx->buffer(5, (size_t)rows * kk * sizeof(double2), &dU));
  QTN_TRY(ctx->buffer(6, (size_t)kk * cols * sizeof(double2), &dV));
  int64_t k = 0;
  QTN_TRY(svd_device(ctx, M, rows, cols, rows, true, er, 0, 0, 1,
                     (double2*)dU, nullptr, (double2*)dV, &k, nullptr));
  // out = U_k (S V^H)_k
  return launch_zgemm(ctx->stream, 0, 0, rows, cols, k,
                      make_double2(1.0, 0.0), (const double2*)dU, rows, 0,
                      (const double2*)dV, k, 0, make_double2(0.0, 0.0), out,
                      rows, 0, 1);
}

}  // namespace
}  // namespace qtn

extern "C" int32_t qtn_contract_svd(qtn_ctx* ctx, const double* A,
                                    int32_t ndA, const int64_t* dimsA,
                                    const double* B, int32_t ndB,
                                    const int64_t* dimsB, int32_t legA,
                                    int32_t legB, double er, double* C) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr && A != nullptr && B != nullptr && C != nullptr &&
              dimsA != nullptr && dimsB != nullptr,
          "qtn_contract_svd: NULL argument");
  QTN_ARG(er >= 0.0, "Error must be positive");
  QTN_ARG(ndA >= 1 && ndA <= 16 && ndB >= 1 && ndB <= 16,
          "qtn_contract_svd: rank must be 1..16");
  QTN_ARG(legA >= 1 && legA <= ndA && legB >= 1 && legB <= ndB,
          "qtn_contract_svd: leg index out of range");
  QTN_ARG(dimsA[legA - 1] == dimsB[legB - 1],
          "Dimensions of contraction legs do not match");
  int64_t nA = 1, nB = 1;
  for (int i = 0; i < ndA; ++i) {
    QTN_ARG(dimsA[i] >= 1, "qtn_contract_svd: empty dimension");
    nA *= dimsA[i];
  }
  for (int i = 0; i < ndB; ++i) {
    QTN_ARG(dimsB[i] >= 1, "qtn_contract_svd: empty dimension");
    nB *= dimsB[i];
  }
  const int64_t D = dimsA[legA - 1];
  const int64_t m1 = nA / D, n2 = nB / D;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  void *dA, *dB, *p1, *p2, *t1, *t2, *dC;
  QTN_TRY(ctx->buffer(8, nA * sizeof(double2), &dA));
  QTN_TRY(ctx->buffer(9, nB * sizeof(double2), &dB));
  QTN_TRY(ctx->buffer(3, nA * sizeof(double2), &p1));
  QTN_TRY(ctx->buffer(4, nB * sizeof(double2), &p2));
  QTN_TRY(ctx->buffer(7, nA * sizeof(double2), &t1));
  QTN_TRY(ctx->buffer(0, nB * sizeof(double2), &t2));
  QTN_TRY(ctx->buffer(10, (size_t)m1 * n2 * sizeof(double2), &dC));
  QTN_CHECK_HIP(hipMemcpyAsync(dA, A, nA * sizeof(double2),
                               hipMemcpyHostToDevice, s));
  QTN_CHECK_HIP(hipMemcpyAsync(dB, B, nB * sizeof(double2),
                               hipMemcpyHostToDevice, s));
  // T1p = permutedims(A, [others..., legA]) as (m1 x D)
  const double2* M1 = (const double2*)dA;
  if (legA != ndA) {
    std::vector<int> perm;
    for (int i = 0; i < ndA; ++i)
      if (i != legA - 1) perm.push_back(i);
    perm.push_back(legA - 1);
    QTN_TRY(qtn::launch_permute(s, (const double2*)dA, ndA, dimsA,
                                perm.data(), (double2*)p1));
    M1 = (const double2*)p1;
  }
  // T2p = permutedims(B, [legB, others...]) as (D x n2)
  const double2* M2 = (const double2*)dB;
  if (legB != 1) {
    std::vector<int> perm;
    perm.push_back(legB - 1);
    for (int i = 0; i < ndB; ++i)
      if (i != legB - 1) perm.push_back(i);
    QTN_TRY(qtn::launch_permute(s, (const double2*)dB, ndB, dimsB,
                                perm.data(), (double2*)p2));
    M2 = (const double2*)p2;
  }
  QTN_TRY(qtn::trunc_reconstruct(ctx, M1, m1, D, er, (double2*)t1));
  QTN_TRY(qtn::trunc_reconstruct(ctx, M2, D, n2, er, (double2*)t2));
  QTN_TRY(qtn::launch_zgemm(s, 0, 0, m1, n2, D, make_double2(1.0, 0.0),
                            (const double2*)t1, m1, 0, (const double2*)t2, D,
                            0, make_double2(0.0, 0.0), (double2*)dC, m1, 0,
                            1));
  QTN_CHECK_HIP(hipMemcpyAsync(C, dC, (size_t)m1 * n2 * sizeof(double2),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}
