// Matrix-product operators on the device: MPO storage, <psi|W|psi> by an
// environment sweep, the compressed application W|psi> by zip-up followed
// by a right-to-left SVD recompression, and the ingest of arbitrary MPS
// site tensors (qtn_mps_set_sites) into the device MPS gauge.
//
// The reference only wires an MPO into a network (apply_MPO,
// src/mpo.jl:184-252) and contracts the whole network with contract(); a
// 64-site chain at chi = 1024 cannot be contracted that way, so the
// compressed application follows the zip-up of Stoudenmire & White,
// New J. Phys. 12, 055026 (2010).
//
// Layouts (column-major, complex FP64):
//   MPS site j   (chi_j, 2, chi_{j+1})           0-based sites, bonds 0..n
//   MPO site j   (D_j, 2 out, 2 in, D_{j+1})     D_0 = D_n = 1
//   environment  L (chi, D, chi') with the bra index first
//   zip-up carry C (k, D, chi): new bond, MPO bond, old bond
#include <algorithm>
#include <cmath>
#include <vector>

#include "qtn_internal.hpp"
#include "qtn_svd.hpp"

struct qtn_mpo {
  qtn_ctx* ctx = nullptr;
  int n = 0;
  std::vector<int64_t> D;      // D_0..D_n
  std::vector<double2*> site;  // W_j
  void* arena = nullptr;
};

namespace qtn {
namespace {

constexpr int kSlotWork = 90, kSlotT1 = 91, kSlotT2 = 92, kSlotU = 93,
              kSlotS = 94, kSlotVh = 95, kSlotScal = 96, kSlotOut = 97,
              kSlotC = 98, kSlotWp = 99;

unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256,
                                                          8192));
}

// out (k x rows, ld k) = V^H for V rows x k (ld rows).
__global__ void adjoint_into(const double2* __restrict__ V, int64_t rows,
                             int64_t k, double2* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       i < rows * k; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = i % k, r = i / k;
    const double2 v = V[r + j * rows];
    out[i] = make_double2(v.x, -v.y);
  }
}

// out[0] = sqrt(sum |A|^2) over n complex values (one workgroup).
__global__ __launch_bounds__(256) void norm2_kernel(
    const double2* __restrict__ A, int64_t n, double* __restrict__ out) {
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const double2 v = A[i];
    s += v.x * v.x + v.y * v.y;
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sqrt(red[0]);
}

// A *= 1 / nrm[0] over n values.
__global__ void scale_by_inv(double2* __restrict__ A, int64_t n,
                             const double* __restrict__ nrm) {
  const double f = nrm[0] > 0.0 ? 1.0 / nrm[0] : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    A[i] = make_double2(A[i].x * f, A[i].y * f);
}

// lam[j] = S[j] / |S| (sum of squares from the smallest value up).
__global__ void normalize_lambda(const double* __restrict__ S, int64_t k,
                                 double* __restrict__ lam) {
  __shared__ double nrm;
  if (threadIdx.x == 0) {
    double w = 0.0;
    for (int64_t j = k - 1; j >= 0; --j) w += S[j] * S[j];
    nrm = sqrt(w);
  }
  __syncthreads();
  const double f = nrm > 0.0 ? 1.0 / nrm : 0.0;
  for (int64_t j = threadIdx.x; j < k; j += blockDim.x) lam[j] = S[j] * f;
}

int32_t permute(hipStream_t s, const double2* in, std::vector<int64_t> dims,
                std::vector<int> perm, double2* out) {
  return launch_permute(s, in, (int)dims.size(), dims.data(), perm.data(),
                        out);
}

int32_t gemm(hipStream_t s, int opA, int opB, int64_t m, int64_t n,
             int64_t k, const double2* A, int64_t lda, const double2* B,
             int64_t ldb, double2* C, int64_t ldc) {
  return launch_zgemm(s, opA, opB, m, n, k, make_double2(1.0, 0.0), A, lda, 0,
                      B, ldb, 0, make_double2(0.0, 0.0), C, ldc, 0, 1);
}

// Right-to-left SVD sweep over left-orthonormal work tensors A_0..A_{n-1}
// (A_j at work[j], dims wd[j], wd[j+1]); the results go to the device MPS
// m: B_j = V_k^H (right-canonical), lambda_j = S_k / |S_k|, and B_0 the
// normalised remainder.  Truncation per bond: tail rule with the relative
// cutoff, then k <= min(chimax, m->maxb[j]).  *norm receives |psi|.
int32_t right_sweep(qtn_mps* m, std::vector<double2*>& work,
                    std::vector<int64_t>& wd, double cutoff, int64_t chimax,
                    double* norm) {
  qtn_ctx* ctx = m->ctx;
  hipStream_t s = ctx->stream;
  const int n = m->n;
  int64_t maxrow = 1, maxcol = 1;
  for (int j = 0; j < n; ++j) {
    maxrow = std::max(maxrow, wd[j]);
    maxcol = std::max(maxcol, 2 * wd[j + 1]);
  }
  const int64_t kk = std::min(maxrow, maxcol);
  void *dU, *dS, *dVh, *dT, *dN;
  QTN_TRY(ctx->buffer(kSlotU, (size_t)maxrow * kk * sizeof(double2), &dU));
  QTN_TRY(ctx->buffer(kSlotS, (size_t)kk * sizeof(double), &dS));
  QTN_TRY(ctx->buffer(kSlotVh, (size_t)kk * maxcol * sizeof(double2), &dVh));
  QTN_TRY(ctx->buffer(kSlotScal, 2 * sizeof(double), &dN));
  std::vector<int64_t> nd(n + 1, 1);
  nd[n] = 1;
  for (int j = n - 1; j >= 1; --j) {
    const int64_t a = wd[j], b = nd[j + 1];
    // A_j as a x (2 b): its right vectors V_k (2b x k) only, then
    // U_k S_k = A_j V_k by a GEMM
    int64_t k = 0;
    const int64_t cap = std::min<int64_t>(chimax > 0 ? chimax : m->cap,
                                          m->maxb[j]);
    QTN_TRY(svd_device_left(ctx, work[j], a, 2 * b, a, 1, cutoff, 1, cap,
                            (double2*)dVh, (double*)dS, &k, nullptr));
    // B_j = V_k^H (k x 2b) straight into the arena (ld k)
    hipLaunchKernelGGL(adjoint_into, dim3(grid_for(2 * b * k)), dim3(256), 0,
                       s, (const double2*)dVh, 2 * b, k, m->site[j]);
    hipLaunchKernelGGL(normalize_lambda, dim3(1), dim3(256), 0, s,
                       (const double*)dS, k, m->lam[j]);
    QTN_CHECK_HIP(hipGetLastError());
    // A_{j-1} <- A_{j-1} (U_k S_k): (2 wd[j-1]) x a times a x k
    QTN_TRY(gemm(s, 0, 0, a, k, 2 * b, work[j], a, (const double2*)dVh,
                 2 * b, (double2*)dU, a));
    const int64_t pa = wd[j - 1];
    QTN_TRY(ctx->buffer(kSlotT1, (size_t)2 * pa * k * sizeof(double2), &dT));
    QTN_TRY(gemm(s, 0, 0, 2 * pa, k, a, work[j - 1], 2 * pa,
                 (const double2*)dU, a, (double2*)dT, 2 * pa));
    QTN_CHECK_HIP(hipMemcpyAsync(work[j - 1], dT,
                                 (size_t)2 * pa * k * sizeof(double2),
                                 hipMemcpyDeviceToDevice, s));
    nd[j] = k;
  }
  // B_0 = A_0 / |A_0|
  const int64_t n0 = 2 * nd[1];
  hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(256), 0, s,
                     (const double2*)work[0], n0, (double*)dN);
  hipLaunchKernelGGL(scale_by_inv, dim3(grid_for(n0)), dim3(256), 0, s,
                     work[0], n0, (const double*)dN);
  QTN_CHECK_HIP(hipGetLastError());
  QTN_CHECK_HIP(hipMemcpyAsync(m->site[0], work[0], n0 * sizeof(double2),
                               hipMemcpyDeviceToDevice, s));
  double nv = 0.0;
  QTN_CHECK_HIP(hipMemcpyAsync(&nv, dN, sizeof(double),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  for (int j = 0; j <= n; ++j) m->dims[j] = nd[j];
  if (norm) *norm = nv;
  return QTN_OK;
}

}  // namespace
}  // namespace qtn

using namespace qtn;

extern "C" {

int32_t qtn_mpo_create(qtn_ctx* ctx, int32_t nsites, const int64_t* D,
                       const double* const* sites, qtn_mpo** out) try {
  clear_error();
  QTN_ARG(ctx && D && sites && out, "qtn_mpo_create: NULL argument");
  QTN_ARG(nsites >= 1, "qtn_mpo_create: need at least one site");
  QTN_ARG(D[0] == 1 && D[nsites] == 1,
          "qtn_mpo_create: boundary bond dimensions must be 1");
  for (int j = 0; j <= nsites; ++j)
    QTN_ARG(D[j] >= 1 && D[j] <= 4096, "qtn_mpo_create: bad bond dimension");
  for (int j = 0; j < nsites; ++j)
    QTN_ARG(sites[j] != nullptr, "qtn_mpo_create: NULL site tensor");
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  qtn_mpo* m = new qtn_mpo();
  m->ctx = ctx;
  m->n = nsites;
  m->D.assign(D, D + nsites + 1);
  size_t bytes = 0;
  std::vector<size_t> off(nsites);
  for (int j = 0; j < nsites; ++j) {
    off[j] = bytes;
    bytes += (size_t)D[j] * 4 * D[j + 1] * sizeof(double2);
    bytes = (bytes + 255) & ~(size_t)255;
  }
  hipError_t e = hipMalloc(&m->arena, bytes);
  if (e != hipSuccess) {
    delete m;
    set_error(std::string("qtn_mpo_create: hipMalloc failed: ") +
              hipGetErrorString(e));
    return QTN_ERR_OOM;
  }
  m->site.resize(nsites);
  for (int j = 0; j < nsites; ++j) {
    m->site[j] = (double2*)((char*)m->arena + off[j]);
    e = hipMemcpyAsync(m->site[j], sites[j],
                       (size_t)D[j] * 4 * D[j + 1] * sizeof(double2),
                       hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
      hipFree(m->arena);
      delete m;
      set_error(std::string("qtn_mpo_create: upload failed: ") +
                hipGetErrorString(e));
      return QTN_ERR_HIP;
    }
  }
  QTN_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  *out = m;
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mpo_destroy(qtn_mpo* m) try {
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

int32_t qtn_mps_expect_mpo(const qtn_mps* m, const qtn_mpo* w, double* out) try {
  return qtn_mps_expect_mpo_form(m, w, QTN_FORM_PLAIN, out);
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_expect_mpo_form(const qtn_mps* m, const qtn_mpo* w,
                                int32_t form, double* out) try {
  clear_error();
  QTN_ARG(m && w && out, "qtn_mps_expect_mpo: NULL argument");
  QTN_ARG(m->n == w->n, "qtn_mps_expect_mpo: MPS and MPO lengths differ");
  QTN_ARG(m->ctx == w->ctx, "qtn_mps_expect_mpo: different contexts");
  QTN_ARG(form == QTN_FORM_PLAIN || form == QTN_FORM_CHOSEN,
          "qtn_mps_expect_mpo: form must be QTN_FORM_PLAIN or "
          "QTN_FORM_CHOSEN");
  // chosen form: the two chi-sized contractions per site run as BF16x3
  // split-precision GEMMs once every dimension reaches 512
  const bool split = form == QTN_FORM_CHOSEN;
  auto big_gemm = [&](hipStream_t st, int opA, int64_t mm, int64_t nn,
                      int64_t kk, const double2* A, int64_t lda,
                      const double2* B, int64_t ldb, double2* C,
                      int64_t ldc) -> int32_t {
    if (split && mm >= 512 && nn >= 512 && kk >= 512)
      return launch_zgemm_bf16x3(m->ctx, opA, mm, nn, kk, A, lda, B, ldb, C,
                                 ldc);
    return gemm(st, opA, 0, mm, nn, kk, A, lda, B, ldb, C, ldc);
  };
  qtn_ctx* ctx = m->ctx;
  hipStream_t s = ctx->stream;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  int64_t chimax = 1, Dmax = 1;
  for (int j = 0; j <= m->n; ++j) {
    chimax = std::max(chimax, m->dims[j]);
    Dmax = std::max(Dmax, w->D[j]);
  }
  const size_t big = (size_t)chimax * chimax * 2 * Dmax;
  void *dL, *dT1, *dT2, *dWp;
  QTN_TRY(ctx->buffer(kSlotWork, big * sizeof(double2), &dL));
  QTN_TRY(ctx->buffer(kSlotT1, big * sizeof(double2), &dT1));
  QTN_TRY(ctx->buffer(kSlotT2, big * sizeof(double2), &dT2));
  QTN_TRY(ctx->buffer(kSlotWp, (size_t)4 * Dmax * Dmax * sizeof(double2),
                      &dWp));
  const double2 one = make_double2(1.0, 0.0);
  QTN_CHECK_HIP(hipMemcpyAsync(dL, &one, sizeof(double2),
                               hipMemcpyHostToDevice, s));
  for (int j = 0; j < m->n; ++j) {
    const int64_t cl = m->dims[j], cr = m->dims[j + 1];
    const int64_t Dl = w->D[j], Dr = w->D[j + 1];
    double2* L = (double2*)dL;
    double2* T1 = (double2*)dT1;
    double2* T2 = (double2*)dT2;
    // T1[a, d, s, b'] = sum_a' L[a, d, a'] B[a', s, b']
    QTN_TRY(big_gemm(s, 0, cl * Dl, 2 * cr, cl, L, cl * Dl, m->site[j], cl,
                     T1, cl * Dl));
    // T2 = T1 as (a, b', d, s)
    QTN_TRY(permute(s, T1, {cl, Dl, 2, cr}, {0, 3, 1, 2}, T2));
    // Wp = W as (d, s, t, e)
    QTN_TRY(permute(s, w->site[j], {Dl, 2, 2, Dr}, {0, 2, 1, 3},
                    (double2*)dWp));
    // T1[a, b', t, e] = sum_{d,s} T2[a, b', d, s] Wp[d, s, t, e]
    QTN_TRY(gemm(s, 0, 0, cl * cr, 2 * Dr, 2 * Dl, T2, cl * cr,
                 (const double2*)dWp, 2 * Dl, T1, cl * cr));
    // T2 = T1 as (a, t, e, b')
    QTN_TRY(permute(s, T1, {cl, cr, 2, Dr}, {0, 2, 3, 1}, T2));
    // L'[b, e, b'] = sum_{a,t} conj(B[a, t, b]) T2[a, t, e, b']
    QTN_TRY(big_gemm(s, 2, cr, Dr * cr, 2 * cl, m->site[j], 2 * cl, T2,
                     2 * cl, L, cr));
  }
  QTN_CHECK_HIP(hipMemcpyAsync(out, dL, sizeof(double2),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_set_sites(qtn_mps* m, const int64_t* dims,
                          const double* const* sites) try {
  clear_error();
  QTN_ARG(m && dims && sites, "qtn_mps_set_sites: NULL argument");
  const int n = m->n;
  QTN_ARG(dims[0] == 1 && dims[3 * (n - 1) + 2] == 1,
          "qtn_mps_set_sites: boundary bond dimensions must be 1");
  for (int j = 0; j < n; ++j) {
    QTN_ARG(sites[j] != nullptr, "qtn_mps_set_sites: NULL site tensor");
    QTN_ARG(dims[3 * j + 1] == 2, "qtn_mps_set_sites: physical dim must be 2");
    QTN_ARG(dims[3 * j] >= 1 && dims[3 * j + 2] >= 1,
            "qtn_mps_set_sites: bad bond dimension");
    if (j + 1 < n)
      QTN_ARG(dims[3 * j + 2] == dims[3 * (j + 1)],
              "Tensor objects last leg must contract with first leg of next "
              "Tensor object");
  }
  qtn_ctx* ctx = m->ctx;
  hipStream_t s = ctx->stream;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  std::vector<int64_t> wd(n + 1);
  size_t total = 0;
  std::vector<size_t> off(n);
  for (int j = 0; j < n; ++j) {
    wd[j] = dims[3 * j];
    off[j] = total;
    total += (size_t)dims[3 * j] * 2 * dims[3 * j + 2];
  }
  wd[n] = 1;
  void* base;
  QTN_TRY(ctx->buffer(kSlotWork, total * sizeof(double2), &base));
  std::vector<double2*> work(n);
  for (int j = 0; j < n; ++j) {
    work[j] = (double2*)base + off[j];
    QTN_CHECK_HIP(hipMemcpyAsync(work[j], sites[j],
                                 (size_t)wd[j] * 2 * wd[j + 1] *
                                     sizeof(double2),
                                 hipMemcpyHostToDevice, s));
  }
  // left sweep: A_j = U (2 wd[j] x k), A_{j+1} <- S V^H A_{j+1}
  int64_t maxrow = 1, maxcol = 1;
  for (int j = 0; j < n; ++j) {
    maxrow = std::max(maxrow, 2 * wd[j]);
    maxcol = std::max(maxcol, wd[j + 1]);
  }
  void *dU, *dS, *dVh, *dT;
  QTN_TRY(ctx->buffer(kSlotU, (size_t)maxrow * maxcol * sizeof(double2),
                      &dU));
  QTN_TRY(ctx->buffer(kSlotS, (size_t)maxcol * sizeof(double), &dS));
  QTN_TRY(ctx->buffer(kSlotVh, (size_t)maxcol * maxcol * sizeof(double2),
                      &dVh));
  for (int j = 0; j + 1 < n; ++j) {
    const int64_t r = 2 * wd[j], c = wd[j + 1];
    int64_t k = 0;
    QTN_TRY(svd_device(ctx, work[j], r, c, r, false, 0.0, 0, 0, 1,
                       (double2*)dU, (double*)dS, (double2*)dVh, &k,
                       nullptr));
    QTN_CHECK_HIP(hipMemcpyAsync(work[j], dU, (size_t)r * k * sizeof(double2),
                                 hipMemcpyDeviceToDevice, s));
    const int64_t nc = 2 * wd[j + 2];
    QTN_TRY(ctx->buffer(kSlotT1, (size_t)k * nc * sizeof(double2), &dT));
    QTN_TRY(gemm(s, 0, 0, k, nc, c, (const double2*)dVh, k, work[j + 1], c,
                 (double2*)dT, k));
    QTN_CHECK_HIP(hipMemcpyAsync(work[j + 1], dT,
                                 (size_t)k * nc * sizeof(double2),
                                 hipMemcpyDeviceToDevice, s));
    wd[j + 1] = k;
  }
  return right_sweep(m, work, wd, 0.0, m->cap, nullptr);
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_apply_mpo(const qtn_mps* in, const qtn_mpo* w, double cutoff,
                          int64_t chimax, qtn_mps* out, double* norm) try {
  return qtn_mps_apply_mpo_form(in, w, cutoff, chimax, QTN_FORM_PLAIN, out,
                                norm);
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_apply_mpo_form(const qtn_mps* in, const qtn_mpo* w,
                               double cutoff, int64_t chimax, int32_t form,
                               qtn_mps* out, double* norm) try {
  clear_error();
  QTN_ARG(in && w && out, "qtn_mps_apply_mpo: NULL argument");
  QTN_ARG(form == QTN_FORM_PLAIN || form == QTN_FORM_CHOSEN,
          "qtn_mps_apply_mpo: form must be QTN_FORM_PLAIN or QTN_FORM_CHOSEN");
  QTN_ARG(in->n == w->n && out->n == in->n,
          "qtn_mps_apply_mpo: lengths differ");
  QTN_ARG(in->ctx == w->ctx && out->ctx == in->ctx,
          "qtn_mps_apply_mpo: different contexts");
  QTN_ARG(in != out, "qtn_mps_apply_mpo: input and output must differ");
  QTN_ARG(cutoff >= 0.0, "Error must be positive");
  QTN_ARG(chimax <= out->cap, "qtn_mps_apply_mpo: chimax above capacity");
  const int64_t cm = chimax > 0 ? chimax : out->cap;
  qtn_ctx* ctx = in->ctx;
  hipStream_t s = ctx->stream;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  const int n = in->n;
  // zip-up output work tensors: bond j at most min(cm, 2^j) (left side)
  std::vector<int64_t> cap(n + 1);
  for (int j = 0; j <= n; ++j)
    cap[j] = std::min<int64_t>(cm, j >= 62 ? cm : ((int64_t)1 << j));
  cap[n] = 1;
  size_t total = 0;
  std::vector<size_t> off(n);
  for (int j = 0; j < n; ++j) {
    off[j] = total;
    total += (size_t)cap[j] * 2 * std::max<int64_t>(cap[j + 1], 1);
  }
  int64_t chi = 1, Dmax = 1;
  for (int j = 0; j <= n; ++j) {
    chi = std::max(chi, in->dims[j]);
    Dmax = std::max(Dmax, w->D[j]);
  }
  void *base, *dC, *dT1, *dT2, *dWp, *dU, *dS, *dVh;
  QTN_TRY(ctx->buffer(kSlotOut, total * sizeof(double2), &base));
  const size_t tsz = (size_t)cm * Dmax * 2 * chi * 2;  // k D s b (+ e)
  QTN_TRY(ctx->buffer(kSlotC, (size_t)cm * Dmax * chi * sizeof(double2),
                      &dC));
  QTN_TRY(ctx->buffer(kSlotT1, tsz * Dmax * sizeof(double2), &dT1));
  QTN_TRY(ctx->buffer(kSlotT2, tsz * Dmax * sizeof(double2), &dT2));
  QTN_TRY(ctx->buffer(kSlotWp, (size_t)4 * Dmax * Dmax * sizeof(double2),
                      &dWp));
  const int64_t kk = std::min<int64_t>(2 * cm, chi * Dmax);
  QTN_TRY(ctx->buffer(kSlotU, (size_t)2 * cm * kk * sizeof(double2), &dU));
  QTN_TRY(ctx->buffer(kSlotS, (size_t)kk * sizeof(double), &dS));
  QTN_TRY(ctx->buffer(kSlotVh, (size_t)kk * chi * Dmax * sizeof(double2),
                      &dVh));
  std::vector<double2*> work(n);
  for (int j = 0; j < n; ++j) work[j] = (double2*)base + off[j];
  std::vector<int64_t> wd(n + 1, 1);
  const double2 one = make_double2(1.0, 0.0);
  QTN_CHECK_HIP(hipMemcpyAsync(dC, &one, sizeof(double2),
                               hipMemcpyHostToDevice, s));
  // The chosen form runs the bond-sized contractions (every dimension at
  // least 512) as BF16x3 split-precision GEMMs; the others, and all of the
  // plain form, in complex FP64.
  const bool split = form == QTN_FORM_CHOSEN;
  auto zgemm = [&](int opA, int64_t mm, int64_t nn, int64_t kk,
                   const double2* A, int64_t lda, const double2* B,
                   int64_t ldb, double2* Cm, int64_t ldc) -> int32_t {
    if (split && mm >= 512 && nn >= 512 && kk >= 512)
      return launch_zgemm_bf16x3(ctx, opA, mm, nn, kk, A, lda, B, ldb, Cm,
                                 ldc);
    return gemm(s, opA, 0, mm, nn, kk, A, lda, B, ldb, Cm, ldc);
  };
  int64_t k = 1;
  bool chosen_off = false;
  double2* C = (double2*)dC;
  double2* T1 = (double2*)dT1;
  double2* T2 = (double2*)dT2;
  for (int j = 0; j < n; ++j) {
    const int64_t cl = in->dims[j], cr = in->dims[j + 1];
    const int64_t Dl = w->D[j], Dr = w->D[j + 1];
    // T1[k, d, s, b] = sum_x C[k, d, x] B[x, s, b]
    QTN_TRY(zgemm(0, k * Dl, 2 * cr, cl, C, k * Dl, in->site[j], cl, T1,
                  k * Dl));
    // T2 = T1 as (k, b, d, s)
    QTN_TRY(permute(s, T1, {k, Dl, 2, cr}, {0, 3, 1, 2}, T2));
    QTN_TRY(permute(s, w->site[j], {Dl, 2, 2, Dr}, {0, 2, 1, 3},
                    (double2*)dWp));
    // T1[k, b, t, e] = sum_{d,s} T2[k, b, d, s] Wp[d, s, t, e]
    QTN_TRY(gemm(s, 0, 0, k * cr, 2 * Dr, 2 * Dl, T2, k * cr,
                 (const double2*)dWp, 2 * Dl, T1, k * cr));
    // T2 = T1 as (k, t, b, e): the (2k) x (cr Dr) matrix of the split
    QTN_TRY(permute(s, T1, {k, cr, 2, Dr}, {0, 2, 1, 3}, T2));
    if (j == n - 1) {
      // cr = Dr = 1: the last site carries the norm
      QTN_CHECK_HIP(hipMemcpyAsync(work[j], T2,
                                   (size_t)2 * k * sizeof(double2),
                                   hipMemcpyDeviceToDevice, s));
      wd[j] = k;
      wd[n] = 1;
      break;
    }
    int64_t kn = 0;
    const int64_t kcap = std::min(cm, cap[j + 1]);
    // The chosen form tries FP32 sweeps on the splits the cap decides
    // until one fails its a-posteriori check (a spectrum flat past the
    // cap); the remaining splits of the zip-up then go straight to FP64, as
    // the following bonds of such a state have equally flat spectra.
    bool fp64 = form != QTN_FORM_CHOSEN || chosen_off;
    if (!fp64) {
      QTN_TRY(svd_device_chosen(ctx, T2, 2 * k, cr * Dr, 2 * k, cutoff, 1,
                                kcap, 1, (double2*)dU, (double*)dS,
                                (double2*)dVh, &kn, nullptr, &fp64));
      if (!fp64)
        QTN_CHECK_HIP(hipMemcpyAsync(work[j], dU,
                                     (size_t)2 * k * kn * sizeof(double2),
                                     hipMemcpyDeviceToDevice, s));
      else if (std::min(2 * k, cr * Dr) > kcap)
        chosen_off = true;
    }
    if (fp64) {
      // U straight into the work tensor (left vectors only: no rotations
      // accumulated, and a wide split's sweeps run on its square R^H),
      // then S V^H = U^H T2
      QTN_TRY(svd_device_left(ctx, T2, 2 * k, cr * Dr, 2 * k, 0, cutoff, 1,
                              kcap, work[j], (double*)dS, &kn, nullptr));
      QTN_TRY(zgemm(2, kn, cr * Dr, 2 * k, work[j], 2 * k, T2, 2 * k,
                    (double2*)dVh, kn));
    }
    wd[j] = k;
    // C' = (S V^H) as (k', b, e) -> (k', e, b)
    QTN_TRY(permute(s, (const double2*)dVh, {kn, cr, Dr}, {0, 2, 1}, C));
    k = kn;
  }
  return right_sweep(out, work, wd, cutoff, cm, norm);
} catch (...) {
  return qtn::abi_exception();
}

}  // extern "C"
