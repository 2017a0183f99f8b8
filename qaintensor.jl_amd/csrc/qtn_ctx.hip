// Context lifetime, error text and device scratch management for libqtn.
#include <algorithm>
#include <cstring>
#include <new>

#include "qtn_internal.hpp"

namespace qtn {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int32_t abi_exception() noexcept {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return QTN_ERR_OOM;
  } catch (const std::exception& e) {
    set_error(std::string("internal error: ") + e.what());
    return QTN_ERR_HIP;
  } catch (...) {
    set_error("internal error: unknown exception");
    return QTN_ERR_HIP;
  }
}

}  // namespace qtn

int32_t qtn_ctx::buffer(int slot, size_t bytes, void** out) {
  if (slot < 0) {
    qtn::set_error("internal: negative buffer slot");
    return QTN_ERR_ARG;
  }
  if ((size_t)slot >= bufs.size()) bufs.resize(slot + 1);
  qtn::DevBuf& b = bufs[slot];
  if (b.bytes < bytes) {
    if (b.ptr) {
      // Pending kernels may still read the old buffer.
      QTN_CHECK_HIP(hipStreamSynchronize(stream));
      QTN_CHECK_HIP(hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
      b.ptr = nullptr;
      qtn::set_error(std::string("hipMalloc of ") + std::to_string(want) +
                     " bytes failed: " + hipGetErrorString(e));
      return QTN_ERR_OOM;
    }
    b.bytes = want;
  }
  *out = b.ptr;
  return QTN_OK;
}

int32_t qtn_ctx::pinned_buffer(size_t bytes, void** out) {
  if (pinned_bytes < bytes) {
    if (pinned) {
      QTN_CHECK_HIP(hipStreamSynchronize(stream));
      QTN_CHECK_HIP(hipHostFree(pinned));
      pinned = nullptr;
      pinned_bytes = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    QTN_CHECK_HIP(hipHostMalloc(&pinned, want, hipHostMallocDefault));
    pinned_bytes = want;
  }
  *out = pinned;
  return QTN_OK;
}

int32_t qtn_ctx::side_stream(hipStream_t* out) {
  if (!side) {
    QTN_CHECK_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    for (auto& e : side_ev)
      QTN_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  *out = side;
  return QTN_OK;
}

int32_t qtn_ctx::sweep_graph(const void* key, size_t n, hipStream_t s,
                             const std::function<int32_t()>& enqueue,
                             hipGraphExec_t* out) {
  for (const auto& g : sweep_graphs)
    if (g.key.size() == n && std::memcmp(g.key.data(), key, n) == 0) {
      *out = g.exec;
      return QTN_OK;
    }
  QTN_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int32_t st = enqueue();
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &graph);
  if (st != QTN_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  QTN_CHECK_HIP(e);
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(graph);
    QTN_CHECK_HIP(ei);
  }
  if (sweep_graphs.size() >= kMaxSweepGraphs) {
    QTN_CHECK_HIP(hipGraphExecDestroy(sweep_graphs.front().exec));
    QTN_CHECK_HIP(hipGraphDestroy(sweep_graphs.front().graph));
    sweep_graphs.erase(sweep_graphs.begin());
  }
  CachedGraph c;
  c.key.assign((const char*)key, (const char*)key + n);
  c.graph = graph;
  c.exec = exec;
  sweep_graphs.push_back(std::move(c));
  *out = exec;
  return QTN_OK;
}

int32_t qtn_ctx::worker(int i, qtn_ctx** out) {
  while ((int)workers.size() <= i) {
    qtn_ctx* w = nullptr;
    QTN_TRY(qtn_ctx_create(device, &w));
    w->parent = this;
    w->prof = prof;
    workers.push_back(w);
  }
  *out = workers[i];
  return QTN_OK;
}

extern "C" {

const char* qtn_last_error(void) { return qtn::g_last_error.c_str(); }

const char* qtn_version(void) { return "libqtn 0.1.0 (gfx950)"; }

int32_t qtn_ctx_create(int32_t device, qtn_ctx** ctx) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr, "qtn_ctx_create: ctx is NULL");
  *ctx = nullptr;
  int ndev = 0;
  QTN_CHECK_HIP(hipGetDeviceCount(&ndev));
  QTN_ARG(device >= 0 && device < ndev,
          "qtn_ctx_create: device " + std::to_string(device) +
              " out of range (" + std::to_string(ndev) + " devices)");
  QTN_CHECK_HIP(hipSetDevice(device));
  QTN_TRY(qtn::noalloc_scratch_init(device));
  qtn_ctx* c = new (std::nothrow) qtn_ctx();
  if (!c) {
    qtn::set_error("qtn_ctx_create: host allocation failed");
    return QTN_ERR_OOM;
  }
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    qtn::set_error(std::string("hipStreamCreate failed: ") +
                   hipGetErrorString(e));
    return QTN_ERR_HIP;
  }
  *ctx = c;
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_ctx_destroy(qtn_ctx* ctx) try {
  qtn::clear_error();
  if (!ctx) return QTN_OK;
  for (qtn_ctx* w : ctx->workers) QTN_TRY(qtn_ctx_destroy(w));
  ctx->workers.clear();
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  QTN_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->side) QTN_CHECK_HIP(hipStreamSynchronize(ctx->side));
  for (auto& b : ctx->bufs)
    if (b.ptr) QTN_CHECK_HIP(hipFree(b.ptr));
  if (ctx->pinned) QTN_CHECK_HIP(hipHostFree(ctx->pinned));
  for (auto e : ctx->ev_pool) QTN_CHECK_HIP(hipEventDestroy(e));
  for (auto& g : ctx->sweep_graphs) {
    QTN_CHECK_HIP(hipGraphExecDestroy(g.exec));
    QTN_CHECK_HIP(hipGraphDestroy(g.graph));
  }
  if (ctx->side) {
    for (auto e : ctx->side_ev) QTN_CHECK_HIP(hipEventDestroy(e));
    QTN_CHECK_HIP(hipStreamDestroy(ctx->side));
  }
  QTN_CHECK_HIP(hipStreamDestroy(ctx->stream));
  delete ctx;
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_ctx_sync(qtn_ctx* ctx) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr, "qtn_ctx_sync: ctx is NULL");
  QTN_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_ctx_counters(qtn_ctx* ctx, int64_t* out, int32_t n,
                         int32_t reset) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr && (n == 0 || out != nullptr),
          "qtn_ctx_counters: NULL argument");
  QTN_ARG(n >= 0 && n <= 6, "qtn_ctx_counters: n must be 0..6");
  // the context's own work plus its workers' (concurrent layer groups)
  int64_t v[6] = {ctx->n_svd, ctx->n_sweeps, ctx->max_sweeps,
                  ctx->n_batches, ctx->n_rr_accepted, ctx->n_rr_redone};
  for (const qtn_ctx* w : ctx->workers) {
    v[0] += w->n_svd;
    v[1] += w->n_sweeps;
    v[2] = std::max(v[2], w->max_sweeps);
    v[3] += w->n_batches;
    v[4] += w->n_rr_accepted;
    v[5] += w->n_rr_redone;
  }
  for (int i = 0; i < n; ++i) out[i] = v[i];
  if (reset) {
    ctx->n_svd = ctx->n_sweeps = ctx->max_sweeps = ctx->n_batches = 0;
    ctx->n_rr_accepted = ctx->n_rr_redone = 0;
    for (qtn_ctx* w : ctx->workers) {
      w->n_svd = w->n_sweeps = w->max_sweeps = w->n_batches = 0;
      w->n_rr_accepted = w->n_rr_redone = 0;
    }
  }
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_ctx_profile(qtn_ctx* ctx, int32_t enable) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr, "qtn_ctx_profile: ctx is NULL");
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  std::vector<qtn_ctx*> all(1, ctx);
  all.insert(all.end(), ctx->workers.begin(), ctx->workers.end());
  for (qtn_ctx* c : all) {
    c->prof = enable != 0;
    for (int k = 0; k < qtn_ctx::kProfKernels; ++k) {
      c->prof_count[k] = 0;
      c->prof_ms[k] = c->prof_bytes[k] = c->prof_flops[k] = 0.0;
    }
  }
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_ctx_profile_read(qtn_ctx* ctx, int32_t kernel, int64_t* count,
                             double* ms, double* bytes, double* flops) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr, "qtn_ctx_profile_read: ctx is NULL");
  QTN_ARG(kernel >= 0 && kernel < qtn_ctx::kProfKernels,
          "qtn_ctx_profile_read: unknown kernel id");
  // the context's own launches plus its workers'
  int64_t c = ctx->prof_count[kernel];
  double m = ctx->prof_ms[kernel], b = ctx->prof_bytes[kernel],
         f = ctx->prof_flops[kernel];
  for (const qtn_ctx* w : ctx->workers) {
    c += w->prof_count[kernel];
    m += w->prof_ms[kernel];
    b += w->prof_bytes[kernel];
    f += w->prof_flops[kernel];
  }
  if (count) *count = c;
  if (ms) *ms = m;
  if (bytes) *bytes = b;
  if (flops) *flops = f;
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_ctx_stream(qtn_ctx* ctx, void** stream) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr && stream != nullptr,
          "qtn_ctx_stream: NULL argument");
  *stream = (void*)ctx->stream;
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

}  // extern "C"
