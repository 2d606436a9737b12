// sd_api.cpp -- the C ABI of libsd_hip.so: contexts, resident matrices, and
// the drivers that replace the reference's R block loops (R/sparseKNN.R,
// R/sparseKNNCross.R, R/sparseSNN.R) with device pipelines.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <thread>

#include <sys/mman.h>
#include <unistd.h>

#include "sd_internal.h"

namespace sd {

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

void clear_error() { g_err.clear(); }

int sync_stream(sd_context* ctx) {
  if (ctx->capturing) return SD_OK;
  SD_HIP(hipStreamSynchronize(ctx->stream));
  return SD_OK;
}

// Blocking copies go through a stream of the library's, never the legacy
// (null) stream: while any thread of the process captures a stream into a
// HIP graph (a server's graph items do), HIP refuses synchronous work on the
// legacy stream in every thread (hipErrorStreamCaptureImplicit).
static int copy_wait(void* dst, const void* src, size_t bytes, hipMemcpyKind kind,
                     hipStream_t s) {
  if (bytes == 0) return SD_OK;
  SD_HIP(hipMemcpyAsync(dst, src, bytes, kind, s));
  SD_HIP(hipStreamSynchronize(s));
  return SD_OK;
}

// One non-blocking stream per device for blocking copies that have no
// context (sd_graph_copy); created on first use, kept until the process ends.
static int transfer_stream(int device, hipStream_t* out) {
  static std::mutex mu;
  static std::vector<hipStream_t> streams;
  std::lock_guard<std::mutex> lk(mu);
  if (device < 0) return set_error(SD_EINVAL, "bad device %d", device);
  if (streams.size() <= (size_t)device) streams.resize((size_t)device + 1, nullptr);
  if (!streams[(size_t)device]) {
    SD_HIP(hipSetDevice(device));
    SD_HIP(hipStreamCreateWithFlags(&streams[(size_t)device], hipStreamNonBlocking));
  }
  *out = streams[(size_t)device];
  return SD_OK;
}

// Maps the pages of a host output array for writing, leaving its contents
// unchanged.  A result array is usually newly allocated and not yet backed
// by memory; the copy that fills it would then take one page fault per 4 KB
// page (about 6 ms for a 24 MB result).  Whole pages inside the array are
// populated by madvise(MADV_POPULATE_WRITE) (Linux 5.14 and later); the
// partial pages at either end, and every page where madvise is refused, by
// writing back the value of one byte of the array that lies in the page.
static void prefault_host(void* ptr, size_t bytes) {
  if (!ptr || bytes == 0) return;
  const long ps = sysconf(_SC_PAGESIZE);
  const uintptr_t page = ps > 0 ? (uintptr_t)ps : 4096u;
  volatile uint8_t* b = static_cast<volatile uint8_t*>(ptr);
  const uintptr_t a0 = (uintptr_t)ptr, a1 = a0 + bytes;
  const uintptr_t p0 = (a0 + page - 1) & ~(page - 1), p1 = a1 & ~(page - 1);
  auto touch = [&](uintptr_t from, uintptr_t to) {  // one byte per page of [from, to)
    for (uintptr_t a = from; a < to; a = (a & ~(page - 1)) + page) {
      volatile uint8_t* q = b + (a - a0);
      *q = *q;
    }
  };
#ifndef MADV_POPULATE_WRITE
#define MADV_POPULATE_WRITE 23
#endif
  if (p1 > p0 && madvise((void*)p0, p1 - p0, MADV_POPULATE_WRITE) == 0) {
    touch(a0, p0);
    touch(p1, a1);
  } else {
    touch(a0, a1);
  }
}

// Prefaults a search's two host result arrays on a thread of its own while
// the device works; wait() (or the destructor) joins it.  Small results are
// left alone (a thread costs more than their faults).
struct HostPrefault {
  std::thread th;
  HostPrefault(void* a, size_t na, void* b, size_t nb) {
    if (na + nb >= ((size_t)1 << 20)) {
      try {
        th = std::thread([=] {
          prefault_host(a, na);
          prefault_host(b, nb);
        });
      } catch (...) {
        // no thread: the copy takes the faults itself
      }
    }
  }
  HostPrefault(const HostPrefault&) = delete;
  HostPrefault& operator=(const HostPrefault&) = delete;
  void wait() {
    if (th.joinable()) th.join();
  }
  ~HostPrefault() { wait(); }
};

// Device-to-host copy of large results into caller arrays (pageable memory):
// one copy into the context's pinned staging buffer, then the host spreads
// it over the arrays on several threads.  A copy straight into pageable
// memory goes through the runtime's own staging in small pieces with one
// CPU copy at a time (about 5 GB/s measured for a 24 MB result).  Falls back
// to direct copies when no staging buffer can be had.
struct HostCopy {
  void* dst;
  const void* src;
  size_t bytes;
};

static int download(sd_context* ctx, const HostCopy* parts, int n) {
  size_t total = 0;
  for (int i = 0; i < n; ++i) total += (parts[i].bytes + 255) & ~(size_t)255;
  if (total >= ((size_t)1 << 20) && ctx->host_stage_bytes < total) {
    if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
    ctx->host_stage = nullptr;
    ctx->host_stage_bytes = 0;
    if (hipHostMalloc(&ctx->host_stage, total, hipHostMallocDefault) == hipSuccess)
      ctx->host_stage_bytes = total;
    else
      ctx->host_stage = nullptr;
    (void)hipGetLastError();
  }
  if (total < ((size_t)1 << 20) || !ctx->host_stage) {
    // the stream is drained before any return: no copy is left writing the
    // caller's arrays
    hipError_t e = hipSuccess;
    for (int i = 0; i < n && e == hipSuccess; ++i)
      if (parts[i].bytes)
        e = hipMemcpyAsync(parts[i].dst, parts[i].src, parts[i].bytes, hipMemcpyDeviceToHost,
                           ctx->stream);
    const hipError_t es = hipStreamSynchronize(ctx->stream);
    SD_HIP(e);
    SD_HIP(es);
    return SD_OK;
  }
  struct Piece {
    void* dst;
    const uint8_t* src;
    size_t bytes;
  };
  std::vector<Piece> pieces;
  uint8_t* stage = static_cast<uint8_t*>(ctx->host_stage);
  size_t off = 0;
  constexpr size_t kPiece = (size_t)1 << 21;
  for (int i = 0; i < n; ++i) {
    if (!parts[i].bytes) continue;
    SD_HIP(hipMemcpyAsync(stage + off, parts[i].src, parts[i].bytes, hipMemcpyDeviceToHost,
                          ctx->stream));
    for (size_t b = 0; b < parts[i].bytes; b += kPiece)
      pieces.push_back({static_cast<uint8_t*>(parts[i].dst) + b, stage + off + b,
                        std::min(kPiece, parts[i].bytes - b)});
    off += (parts[i].bytes + 255) & ~(size_t)255;
  }
  SD_HIP(hipStreamSynchronize(ctx->stream));
  std::atomic<size_t> next{0};
  auto work = [&] {
    for (size_t j; (j = next.fetch_add(1)) < pieces.size();)
      std::memcpy(pieces[j].dst, pieces[j].src, pieces[j].bytes);
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t nt = std::min<size_t>({pieces.size(), 8, std::max(1u, hw / 2)});
  std::vector<std::thread> helpers;
  for (size_t t = 1; t < nt; ++t) {
    try {
      helpers.emplace_back(work);
    } catch (...) {
      break;  // fewer helpers; the calling thread copies the rest
    }
  }
  work();
  for (std::thread& h : helpers) h.join();
  return SD_OK;
}

std::atomic<int64_t> g_device_allocs{0};

int DevBuf::alloc(size_t n) {
  g_device_allocs.fetch_add(1, std::memory_order_relaxed);
  release();
  if (n == 0) n = 16;
  hipError_t e = pool ? hipMallocAsync(&ptr, n, pool) : hipMalloc(&ptr, n);
  if (e != hipSuccess) {
    ptr = nullptr;
    (void)hipGetLastError();
    return set_error(SD_ENOMEM, "device allocation of %zu bytes failed: %s", n,
                     hipGetErrorString(e));
  }
  bytes = n;
  return SD_OK;
}

int DevBuf::ensure(size_t n) {
  if (ptr && bytes >= n) return SD_OK;
  return alloc(n);
}

void DevBuf::release() {
  if (ptr) (void)(pool ? hipFreeAsync(ptr, pool) : hipFree(ptr));
  ptr = nullptr;
  bytes = 0;
}

KernelTimer::KernelTimer(sd_context* c, double* a, int64_t* n)
    : ctx(c), acc(a), count(n), on(c->timing) {
  if (on) (void)hipEventRecord(ctx->ev_a, ctx->stream);
}

KernelTimer::~KernelTimer() {
  if (!on) return;
  (void)hipEventRecord(ctx->ev_b, ctx->stream);
  if (hipEventSynchronize(ctx->ev_b) == hipSuccess) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev_a, ctx->ev_b) == hipSuccess) *acc += ms;
    *count += 1;
  }
}

int context_use_pool(sd_context* ctx) {
  hipMemPool_t pool = nullptr;
  SD_HIP(hipDeviceGetDefaultMemPool(&pool, ctx->device));
  // keep freed blocks in the pool instead of returning them at each
  // synchronisation, so a worker's buffers are reused without new allocations
  uint64_t keep = UINT64_MAX;
  SD_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
  for (auto& b : ctx->scratch) b.pool = ctx->stream;
  for (auto& b : ctx->snn_scratch) b.pool = ctx->stream;
  for (DevBuf* b : {&ctx->xm.thr, &ctx->xm.cnt, &ctx->xm.fkey, &ctx->xm.fidx})
    b->pool = ctx->stream;
  return SD_OK;
}

void matrix_use_pool(sd_matrix* m, hipStream_t stream) {
  for (DevBuf* b : {&m->p, &m->i, &m->x, &m->deg, &m->unit, &m->scale, &m->unit_ssq, &m->rows16,
                    &m->mean, &m->css, &m->cmf, &m->cif, &m->sqn, &m->sqf})
    b->pool = stream;
  for (auto& b : m->dense) b.pool = stream;
}

// Validate a CSC matrix, drop explicit zeros, compute the column flags and
// copy it into m's device buffers (grown only when too small), discarding
// every layout derived from earlier contents.
int matrix_load(sd_context* ctx, sd_matrix* m, int64_t nrow, int64_t ncol, const int64_t* colptr,
                const int32_t* rowind, const double* values) {
  if (!colptr) return set_error(SD_EINVAL, "NULL argument");
  if (nrow < 0 || ncol < 0) return set_error(SD_EINVAL, "negative dimensions");
  if (nrow >= ((int64_t)1 << 31)) return set_error(SD_EINVAL, "nrow must be below 2^31");
  if (colptr[0] != 0) return set_error(SD_EINVAL, "colptr[0] must be 0");
  const int64_t nnz_in = colptr[ncol];
  if (nnz_in > 0 && (!rowind || !values)) return set_error(SD_EINVAL, "NULL argument");
  // validate, drop explicit zeros, flags (one host pass)
  std::vector<int64_t> p((size_t)ncol + 1);
  std::vector<int32_t> ri;
  std::vector<double> vx;
  ri.reserve((size_t)nnz_in);
  vx.reserve((size_t)nnz_in);
  std::vector<uint8_t> hdeg((size_t)ncol, 0);
  bool any_nonfinite = false, any_negative = false;
  double abs_min = INFINITY, abs_max = 0.0;
  p[0] = 0;
  for (int64_t j = 0; j < ncol; ++j) {
    if (colptr[j + 1] < colptr[j]) return set_error(SD_EINVAL, "colptr must be non-decreasing");
    int32_t prev = -1;
    for (int64_t e = colptr[j]; e < colptr[j + 1]; ++e) {
      const int32_t r = rowind[e];
      if (r < 0 || r >= nrow || r <= prev)
        return set_error(SD_EINVAL,
                         "row indices must be in [0, nrow) and strictly ascending per column");
      prev = r;
      const double v = values[e];
      if (v == 0.0) continue;
      if (!std::isfinite(v)) {
        any_nonfinite = true;
      } else {
        abs_min = std::min(abs_min, std::fabs(v));
        abs_max = std::max(abs_max, std::fabs(v));
      }
      if (v < 0.0) any_negative = true;
      ri.push_back(r);
      vx.push_back(v);
    }
    p[(size_t)j + 1] = (int64_t)ri.size();
    // colSums(abs(.)) == 0 <=> no stored non-zero value (NaN is not zero)
    hdeg[(size_t)j] = (p[(size_t)j + 1] == p[(size_t)j]) ? 1 : 0;
  }
  const int64_t nnz = (int64_t)ri.size();
  SD_TRY(m->p.ensure(sizeof(int64_t) * (size_t)(ncol + 1)));
  SD_TRY(m->i.ensure(sizeof(int32_t) * (size_t)(nnz + 1)));
  SD_TRY(m->x.ensure(sizeof(double) * (size_t)(nnz + 1)));
  SD_TRY(m->deg.ensure((size_t)ncol + 16));
  SD_HIP(hipMemcpyAsync(m->p.ptr, p.data(), sizeof(int64_t) * (size_t)(ncol + 1),
                        hipMemcpyHostToDevice, ctx->stream));
  if (nnz) {
    SD_HIP(hipMemcpyAsync(m->i.ptr, ri.data(), sizeof(int32_t) * (size_t)nnz,
                          hipMemcpyHostToDevice, ctx->stream));
    SD_HIP(hipMemcpyAsync(m->x.ptr, vx.data(), sizeof(double) * (size_t)nnz,
                          hipMemcpyHostToDevice, ctx->stream));
  }
  if (ncol)
    SD_HIP(hipMemcpyAsync(m->deg.ptr, hdeg.data(), (size_t)ncol, hipMemcpyHostToDevice,
                          ctx->stream));
  SD_HIP(hipStreamSynchronize(ctx->stream));
  m->ctx = ctx;
  m->device = ctx->device;
  m->nrow = nrow;
  m->ncol = ncol;
  m->nnz = nnz;
  m->hp = std::move(p);
  m->hdeg = std::move(hdeg);
  m->any_nonfinite = any_nonfinite;
  m->any_negative = any_negative;
  m->abs_min = abs_min;
  m->abs_max = abs_max;
  m->unit_ready = m->centre_ready = m->sq_ready = false;
  m->tuned_margin = 0;
  m->raw_split = false;
  for (int t = 0; t < kDenseKinds; ++t) m->dense_kpad[t] = 0;
  return SD_OK;
}

}  // namespace sd

using namespace sd;

namespace {

int check_cancel(sd_context* ctx) {
  if (ctx->cancel.load(std::memory_order_relaxed))
    return set_error(SD_ECANCELLED, "Computation interrupted by the user.");
  return SD_OK;
}

// Ranking direction (R/sparseKNN.R:168-181).
int method_direction(int method, int dist, int* dec) {
  switch (method) {
    case SD_BINARY:
    case SD_COSINE:
    case SD_PEARSON:
      *dec = !dist;
      return SD_OK;
    case SD_COVARIANCE:
      *dec = 1;
      return SD_OK;
    case SD_EUCLIDEAN:
    case SD_MANHATTAN:
    case SD_JS:
      if (!dist)
        return set_error(SD_EINVAL,
                         "'dist = FALSE' is only meaningful for methods \"binary\", \"cosine\" and "
                         "\"pearson\"; %s has no similarity form.",
                         method == SD_EUCLIDEAN ? "euclidean"
                         : method == SD_JS      ? "js"
                                                : "manhattan");
      *dec = 0;
      return SD_OK;
    default:
      return set_error(SD_EUNSUPPORTED, "method %d is not provided by this library", method);
  }
}

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

struct DevOut {
  DevBuf idx, val;
};

// Prepare per-method statistics of a matrix.
int prep_method(sd_context* ctx, sd_matrix* m, int method) {
  if (method == SD_COSINE) {
    if (m->any_nonfinite)
      return set_error(SD_ENONFINITE, "Cosine requires finite input (found NA, NaN or Inf).");
    return prep_unit(ctx, m);
  }
  if (method == SD_PEARSON || method == SD_COVARIANCE) return prep_centre(ctx, m);
  if (method == SD_JS && (m->any_nonfinite || m->any_negative))
    return set_error(SD_EINVAL,
                     "fastJS2 requires finite, non-negative input (columns should be normalised "
                     "to sum to 1 for a true JS divergence).");
  return SD_OK;
}

ExactArgs exact_args(const sd_matrix* ref, const sd_matrix* qry, int method, int dist) {
  ExactArgs a;
  a.method = method;
  a.dist = dist;
  a.ref = ref->view();
  a.qry = qry->view();
  a.ref_unit = ref->unit.as<double>();
  a.ref_scale = ref->scale.as<double>();
  a.ref_rows16 = (ref->unit_ready && matrix_rows16(ref)) ? ref->rows16.as<uint16_t>() : nullptr;
  a.qry_unit = qry->unit.as<double>();
  a.qry_scale = qry->scale.as<double>();
  a.ref_mean = ref->mean.as<double>();
  a.ref_css = ref->css.as<double>();
  a.qry_mean = qry->mean.as<double>();
  a.qry_css = qry->css.as<double>();
  return a;
}

// Exact FP64 slab + top-k for the query columns listed in d_qcols (device),
// writing results into out (nq x k column-major) at those columns.  Memory
// is bounded by processing the columns in slabs.
int exact_knn_columns(sd_context* ctx, const ExactArgs& a, const int32_t* d_qcols, int64_t ncols,
                      int64_t nq_total, int k, int dec, int64_t block_hint, int32_t* d_out_idx,
                      double* d_out_val) {
  const int64_t nref = a.ref.ncol;
  const int64_t budget = (int64_t)4 << 30;  // bytes of slab per pass
  int64_t nb = budget / (int64_t)std::max<int64_t>(1, nref * 8);
  nb = std::max<int64_t>(1, std::min<int64_t>(nb, 65535));
  if (block_hint > 0) nb = std::min<int64_t>(nb, std::max<int64_t>(block_hint, 1));
  nb = std::min<int64_t>(nb, ncols);
  DevBuf& D = ctx->scratch[0];
  SD_TRY(D.ensure(sizeof(double) * (size_t)nref * (size_t)nb + 64));
  for (int64_t c0 = 0; c0 < ncols; c0 += nb) {
    // the host waits for each block before queueing the next, so a cancel
    // request stops the search within one block
    if (c0 > 0) SD_HIP(hipStreamSynchronize(ctx->stream));
    SD_TRY(check_cancel(ctx));
    const int64_t n = std::min(nb, ncols - c0);
    SD_TRY(exact_slab(ctx, a, d_qcols + c0, n, D.as<double>(), nref, 1));
    SD_TRY(topk_dense(ctx, D.as<double>(), nref, n, nref, k, dec, nullptr, d_out_idx, d_out_val,
                      nq_total, d_qcols + c0));
  }
  SD_HIP(hipStreamSynchronize(ctx->stream));
  return SD_OK;
}

struct KnnPlan {
  int k = 0, dec = 0, warn = 0;
  bool selfmode = false;
};

// Validation, degenerate masks and k clamp (R/sparseKNN.R:114-231).
int plan_knn(const sd_matrix* X, const sd_matrix* Y, const sd_knn_params* p, KnnPlan* out,
             int64_t* avail_out) {
  if (p->k < 1)
    return set_error(SD_EINVAL,
                     "'k' must be a positive integer no larger than .Machine$integer.max.");
  if (p->block_size < 0)
    return set_error(SD_EINVAL,
                     "'block_size' must be a positive integer no larger than .Machine$integer.max.");
  SD_TRY(method_direction(p->method, p->dist, &out->dec));
  out->selfmode = (Y == nullptr);
  if (X->ncol < 1) return set_error(SD_EINVAL, "'X' must have at least one column.");
  if (!out->selfmode && X->nrow != Y->nrow)
    return set_error(SD_EINVAL, "'X' and 'Y' must have the same number of rows.");
  const sd_matrix* ref = out->selfmode ? X : Y;
  int64_t avail = 0;
  for (int64_t j = 0; j < ref->ncol; ++j) avail += !(p->drop_degenerate && ref->hdeg[j]);
  if (out->selfmode && !p->include_self) avail -= 1;
  *avail_out = avail;
  if (avail < 1) return set_error(SD_ENOCAND, "No candidate neighbours available.");
  out->k = p->k;
  if ((int64_t)out->k > avail) {
    out->warn = 1;
    out->k = (int)avail;
  }
  return SD_OK;
}

// Per-query candidate lists (top-L by the slab key) for every query of fa.
//
// Large searches run in two passes.  A list-mode pass over every stride-th
// reference tile gives each query a threshold (its L-th key there), a valid
// lower bound for its L-th key over all references.  The filter-mode pass
// then computes every (query tile, reference tile) exactly once in a
// rasterised grid -- concurrent workgroups share their tiles in L2 -- and
// appends only candidates at or above the threshold; a per-query top-L of
// those is the result.  Queries whose buffer overflows (many ties at the
// threshold) get a full list-mode pass instead.  Small searches use the
// list mode directly.
// The filtered pass's candidate buffers, kept after the lists are selected:
// every reference at or above its query's threshold, so the top-L' of a
// buffer that holds at least L' entries and did not overflow is the query's
// top-L' for any L' (fkey is null when the search took list mode only).
struct CandBuffers {
  const double* fkey = nullptr;
  const int32_t* fidx = nullptr;
  const int32_t* cnt = nullptr;
  int cap = 0;
  // a search timed for the sample tuner: its class, candidate and scored
  // pairs, timed from the pre-pass to the end of the search
  bool record = false;
  SampleKey key;
  int cand = 0;
  double pairs = 0;
  int64_t allocs = 0;  // g_device_allocs at the start
  std::chrono::steady_clock::time_point t_start;
};

int fused_candidate_lists(sd_context* ctx, FusedArgs fa, const double** out_key,
                          const int32_t** out_idx, CandBuffers* cb) {
  const int64_t nq = fa.nqry;
  const int L = fa.L;
  const int64_t rtiles = (fa.nref + 127) / 128;
  auto list_pass = [&](FusedArgs a, DevBuf& kbuf, DevBuf& ibuf, DevBuf& mkbuf, DevBuf& mibuf,
                       const double** rk, const int32_t** ri) -> int {
    const int64_t qtiles = (a.nqry + 255) / 256;
    // two workgroups per CU at least, the references split in multiples of
    // 8 (one per XCD); query sets of eight query tiles per CU or more with
    // short rows (2 KB or less) fill the GPU with one split, whose lists then
    // need no merge.  With long rows one split was slower: each workgroup
    // streams every sampled reference tile (7.7 MB each at 30,000 rows).
    int64_t splits = (2 * (int64_t)ctx->cu_count + qtiles - 1) / qtiles;
    const bool one_split =
        qtiles >= 8 * (int64_t)ctx->cu_count && dense_row_bytes(a.kind, a.kpad) <= 2048;
    splits = one_split ? 1 : round_up(std::max<int64_t>(splits, 1), 8);
    a.splits = (int)std::min<int64_t>(splits, 4096);
    const int nlists = fused_effective_splits(a);
    const int La = a.L;
    SD_TRY(kbuf.ensure(sizeof(double) * (size_t)nlists * (size_t)a.nqry * (size_t)La + 64));
    SD_TRY(ibuf.ensure(sizeof(int32_t) * (size_t)nlists * (size_t)a.nqry * (size_t)La + 64));
    SD_TRY(fused_slab_topl(ctx, a, kbuf.as<double>(), ibuf.as<int32_t>()));
    *rk = kbuf.as<double>();
    *ri = ibuf.as<int32_t>();
    if (nlists > 1) {
      SD_TRY(mkbuf.ensure(sizeof(double) * (size_t)a.nqry * (size_t)La + 64));
      SD_TRY(mibuf.ensure(sizeof(int32_t) * (size_t)a.nqry * (size_t)La + 64));
      SD_TRY(merge_lists(ctx, *rk, *ri, nlists, a.nqry, La, mkbuf.as<double>(),
                         mibuf.as<int32_t>()));
      *rk = mkbuf.as<double>();
      *ri = mibuf.as<int32_t>();
    }
    return SD_OK;
  };
  // The pre-pass samples every stride-th reference tile and takes each
  // query's Lp-th key there as its threshold.  About stride x Lp references
  // of the whole set score at or above it (the Lp-th of a 1/stride sample),
  // so a sparser sample with a proportionally smaller rank keeps the
  // candidate count and costs less: 1/128 at rank ~L/4 from 512 tiles on,
  // 1/64 at rank ~L/2 from 256, 1/32 at rank L below (a sample of at least
  // 4 tiles).
  CrossMirror& xm = ctx->xm;
  if (rtiles < 4 * 32) {
    if (xm.phase == 1) {  // probe of a reverse search without a filtered pass
      *out_key = nullptr;
      *out_idx = nullptr;
      return SD_OK;
    }
    fa.tile_stride = 1;
    return list_pass(fa, ctx->scratch[3], ctx->scratch[4], ctx->scratch[5], ctx->scratch[6],
                     out_key, out_idx);
  }
  // Every candidate costs the filtered pass per-tile epilogue work (exact
  // tests, appends) that a long K amortises and a short K does not, while a
  // sampled pair costs the pre-pass K operations.  Short rows (embeddings of
  // a few hundred dimensions) therefore take a denser sample at a low rank:
  // every 64th tile at rank ~L/4 (at least 8) leaves about 8 x 64 = 512
  // candidates per query instead of about 1000.  On 256-dimensional
  // embeddings (200,000 x 1,000,000 cross search) every 32nd, 64th and
  // 128th tile gave 339, 318 and 329 ms.
  // That rule is the default; searches of a class large enough to time
  // choose among the candidates by measured time (sample_plan).
  const int64_t row_bytes = dense_row_bytes(fa.kind, fa.kpad);
  const bool short_rows = row_bytes <= 2048;
  const int def = short_rows ? 3 : rtiles >= 4 * 128 ? 2 : rtiles >= 4 * 64 ? 1 : 0;
  bool allowed[kSampleCandidates];
  for (int c = 0; c < kSampleCandidates; ++c)
    allowed[c] = c == def || rtiles >= 4 * (int64_t)sample_candidate(c, L).stride;
  // the symmetric pass scores each unordered pair once
  const bool tri_pass = fa.self_mode && !fa.qlist && fa.nqry == fa.nref;
  const double pairs = (double)nq * (double)fa.nref * (tri_pass ? 0.5 : 1.0);
  SampleKey skey;
  skey.method = fa.method;
  skey.kind = fa.kind;
  skey.density = fa.ref_density_bucket;
  skey.row_bytes = (int)std::ceil(std::log2((double)std::max<int64_t>(row_bytes, 1)));
  // list lengths move with the re-rank margin; the candidates' ranks scale
  // with L, so short and long lists form the two classes
  skey.L = L <= 16 ? 16 : 64;
  skey.tri = tri_pass ? 1 : 0;
  bool record = false;
  // cross pairs sharing one pass keep the rule (both directions' buffers
  // are sized from it)
  const int cand = sample_plan(skey, allowed, def, ctx->sample_tuning != 0 && xm.phase == 0,
                               pairs >= (double)((int64_t)1 << 28) && !ctx->capturing, &record);
  sample_note(skey, def, cand);
  const SampleChoice sc = sample_candidate(cand, L);
  const int64_t stride = sc.stride;
  const int Lp = sc.rank;
  ctx->stats.sample_stride = stride;
  ctx->stats.sample_rank = Lp;
  cb->record = record;
  cb->key = skey;
  cb->cand = cand;
  cb->pairs = pairs;
  cb->allocs = g_device_allocs.load(std::memory_order_relaxed);
  cb->t_start = std::chrono::steady_clock::now();
  // room for three times the expected candidates (the count above the
  // Lp-th of a sample varies by about 1/sqrt(Lp))
  int cap = (int)std::min<int64_t>(4096, 3 * Lp * stride + 256);
  // a search the cross mirror can serve: same layout, precision and method,
  // all query columns, a key without per-query weights
  auto mirrorable = [&](const FusedArgs& f) {
    return xm.ready && !f.qlist && f.kind == xm.kind && f.kpad == xm.kpad && f.prec == xm.prec &&
           f.method == xm.method;
  };
  DevBuf* cntb = &ctx->scratch[14];
  DevBuf* fkb = &ctx->scratch[9];
  DevBuf* fib = &ctx->scratch[10];
  if (xm.phase == 3 && xm.target.filled && mirrorable(fa) && fa.L == xm.L && fa.nqry == xm.nqry &&
      fa.nref == xm.nref) {
    // the reverse search of a cross pair: its buffers were filled by the
    // forward search's filtered pass
    cntb = &xm.cnt;
    fkb = &xm.fkey;
    fib = &xm.fidx;
    cap = xm.target.cap;
  } else {
    // 1. thresholds from every stride-th reference tile
    const double* pk = nullptr;
    const int32_t* pi = nullptr;
    FusedArgs pre = fa;
    pre.tile_stride = (int)stride;
    pre.L = Lp;
    SD_TRY(list_pass(pre, ctx->scratch[3], ctx->scratch[4], ctx->scratch[5], ctx->scratch[6], &pk,
                     &pi));
    DevBuf& thr = xm.phase == 1 ? xm.thr : ctx->scratch[13];
    SD_TRY(thr.ensure(sizeof(double) * (size_t)nq + 64));
    SD_TRY(list_thresholds(ctx, pk, pi, nq, Lp, thr.as<double>()));
    if (xm.phase == 1) {
      // probe of the reverse search: keep its thresholds and buffers for the
      // forward search's filtered pass
      // the buffers stay allocated while the forward search runs; without
      // room for them the two searches simply run one after the other
      const bool room = !fa.qlist &&
                        xm.cnt.ensure(sizeof(int32_t) * (size_t)(nq + 16)) == SD_OK &&
                        xm.fkey.ensure(sizeof(double) * (size_t)nq * (size_t)cap + 64) == SD_OK &&
                        xm.fidx.ensure(sizeof(int32_t) * (size_t)nq * (size_t)cap + 64) == SD_OK;
      if (!room) {
        xm.cnt.release();
        xm.fkey.release();
        xm.fidx.release();
        clear_error();
      } else {
        xm.L = L;
        xm.kind = fa.kind;
        xm.prec = fa.prec;
        xm.method = fa.method;
        xm.kpad = fa.kpad;
        xm.nqry = nq;
        xm.nref = fa.nref;
        xm.target = MirrorTarget{thr.as<double>(),    cap,        xm.cnt.as<int32_t>(),
                                 xm.fkey.as<double>(), xm.fidx.as<int32_t>(), fa.ref_deg,
                                 false,                fa.dot_e0, fa.dot_e1, fa.dot_nonneg};
        xm.ready = true;
      }
      *out_key = nullptr;
      *out_idx = nullptr;
      return SD_OK;
    }
    // 2. filtered full pass (a cross pair's forward search also feeds the
    // reverse search's buffers)
    SD_TRY(cntb->ensure(sizeof(int32_t) * (size_t)(nq + 16)));
    SD_TRY(fkb->ensure(sizeof(double) * (size_t)nq * (size_t)cap + 64));
    SD_TRY(fib->ensure(sizeof(int32_t) * (size_t)nq * (size_t)cap + 64));
    SD_TRY(check_cancel(ctx));
    FusedArgs ff = fa;
    if (xm.phase == 2 && mirrorable(fa) && xm.nqry == fa.nref && xm.nref == fa.nqry)
      ff.mirror = &xm.target;
    SD_TRY(fused_filter_pass(ctx, ff, thr.as<double>(), cap, cntb->as<int32_t>(),
                             fkb->as<double>(), fib->as<int32_t>()));
  }
  DevBuf& cnt = *cntb;
  DevBuf& fk = *fkb;
  DevBuf& fi = *fib;
  // 3. per-query top-L of the appended candidates
  DevBuf& mk = ctx->scratch[5];
  DevBuf& mi = ctx->scratch[6];
  DevBuf& ov = ctx->scratch[15];
  SD_TRY(mk.ensure(sizeof(double) * (size_t)nq * (size_t)L + 64));
  SD_TRY(mi.ensure(sizeof(int32_t) * (size_t)nq * (size_t)L + 64));
  SD_TRY(ov.ensure(sizeof(int32_t) * (size_t)(nq + 16)));
  int32_t* over_count = ov.as<int32_t>() + nq;
  SD_TRY(zero_i32(ctx, over_count, 1));
  // A buffer holds exactly the references whose filtered-pass key is at or
  // above the threshold.  With at least L of them, the L-th key is >= thr and
  // every reference outside keys below it, so the buffer's top-L is the
  // query's top-L.  That needs nothing from the pre-pass: a threshold from
  // rank Lp < L can exceed the true L-th key, and the filtered pass may
  // compute a key in another order than the pre-pass did (the symmetric
  // pass's mirrored values on the split layout); either way fewer than L
  // candidates send the query to a full pass instead.
  SD_TRY(select_lists(ctx, fk.as<double>(), fi.as<int32_t>(), cnt.as<int32_t>(), cap, nq, L,
                      mk.as<double>(), mi.as<int32_t>(), ov.as<int32_t>(), over_count, L));
  cb->fkey = fk.as<double>();
  cb->fidx = fi.as<int32_t>();
  cb->cnt = cnt.as<int32_t>();
  cb->cap = cap;
  *out_key = mk.as<double>();
  *out_idx = mi.as<int32_t>();
  if (ctx->capturing) {
    // a captured search leaves the count for its caller to read after replay
    ctx->cap_over = over_count;
    return SD_OK;
  }
  int32_t nover = 0;
  SD_HIP(hipMemcpyAsync(&nover, over_count, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  SD_HIP(hipStreamSynchronize(ctx->stream));
  ctx->stats.overflow_queries += nover;
  if (nover > 0) {
    // 4. overflowed queries: full list-mode pass over all reference tiles
    FusedArgs ofa = fa;
    ofa.tile_stride = 1;
    ofa.qlist = ov.as<int32_t>();
    ofa.nqry = nover;
    const double* ok = nullptr;
    const int32_t* oi = nullptr;
    SD_TRY(list_pass(ofa, ctx->scratch[3], ctx->scratch[4], ctx->scratch[11], ctx->scratch[12], &ok,
                     &oi));
    SD_TRY(scatter_lists(ctx, ok, oi, ov.as<int32_t>(), nover, L, mk.as<double>(),
                         mi.as<int32_t>()));
  }
  return SD_OK;
}

// Constants of the euclidean and pearson keys' error bound (euc_key and
// pear_key in k_slab.hip): the slab's dot product x of a query q with
// nnz(q) = m stored values against a reference r obeys
// |x - q.r| <= (e0 + e1 m) sum |q_i r_i|.
//  * split BF16 (hi = bf16(v), lo = bf16(v - hi), each within ub = 2^-8 +
//    2^-23 relative, the second term of ub for the FP32 step): the three
//    products hi.hi + hi.lo + lo.hi miss lo.lo and the rounding of lo, at most
//    3.1 ub^2 |q_i r_i| per row; the FP32 accumulation adds at most 3 m terms
//    whose magnitudes sum to at most 1.02 sum |q_i r_i|, each addition within
//    2^-23 relative (twice the round-to-nearest unit, so the bound holds for
//    any rounding of the accumulator), inflated by 1 / (1 - n 2^-23) for the
//    n = 3 nrow additions of the longest column.
//  * plain BF16: each product within (1 + ub)^2 - 1 = 2 ub + ub^2; m terms.
//  * FP32: each input within 2^-24 and each product rounded once (2^-23
//    allowed): 4 2^-23 per row; m terms.
// For non-negative data the key needs q.r <= x / (1 - e), so both constants
// are divided by 1 - e at the longest column.  The path is taken only for
// finite values in [1e-15, 1e15] (no FP32 overflow of squares or norms, no
// subnormal BF16 parts) and e <= 0.25.
struct DotBound {
  float e0 = 0.f, e1 = 0.f;
  int nonneg = 0;
};
bool dot_bound(const sd_matrix* ref, const sd_matrix* X, int kind, DotBound* out) {
  for (const sd_matrix* m : {ref, X}) {
    if (m->any_nonfinite) return false;
    if (m->nnz > 0 && (m->abs_min < 1e-15 || m->abs_max > 1e15)) return false;
    if ((double)m->nrow * m->abs_max * m->abs_max > 1e36) return false;
  }
  const double nrow = (double)std::max<int64_t>(X->nrow, 1);
  const double u23 = std::ldexp(1.0, -23);
  const double ub = std::ldexp(1.0, -8) + u23;
  const double terms = kind == kDenseBf16Split ? 3.0 : 1.0;  // accumulated terms per row
  const double n = terms * nrow;
  if (n * u23 >= 0.2) return false;
  const double g = 1.0 / (1.0 - n * u23);
  double e0, mag;  // per-row representation error; magnitude of the summed terms
  if (kind == kDenseBf16Split) {
    e0 = 3.1 * ub * ub;
    mag = 1.02;
  } else if (kind == kDenseBf16Raw) {
    e0 = 2.0 * ub + ub * ub;
    mag = (1.0 + ub) * (1.0 + ub);
  } else {
    e0 = 4.0 * u23;
    mag = 1.0 + 4.0 * u23;
  }
  const double e1 = terms * mag * u23 * g;
  const double emax = e0 + e1 * nrow;
  if (emax > 0.25) return false;
  const double fold = 1.0 / (1.0 - emax);
  // round the constants up to FP32
  out->e0 = std::nextafter((float)(e0 * fold), INFINITY);
  out->e1 = std::nextafter((float)(e1 * fold), INFINITY);
  out->nonneg = (!ref->any_negative && !X->any_negative) ? 1 : 0;
  return true;
}

// Query columns per block of a fused search.  Per query the search holds up
// to 4096 filtered-pass candidates, the pre-pass's split lists (at most 8 of
// rank <= L, before and after merging), the selected lists and the wider
// escalation lists, 12 bytes an entry (FP64 key, int32 index); a budget --
// the context's, else half the free device memory -- divided by that gives
// the block, a multiple of the 256-column query tile.
int64_t fused_query_block(sd_context* ctx, int64_t nq, int L) {
  if (ctx->capturing) return nq;  // captured searches are a few query tiles
  int64_t budget = ctx->buffer_budget;
  if (budget <= 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      (void)hipGetLastError();
      return nq;
    }
    budget = (int64_t)(free_b / 2);
  }
  const int64_t Lmax = fused_list_len(64);
  const int64_t per_query = 12 * (4096 + 16 * (int64_t)L + (int64_t)L + 3 * Lmax) + 64;
  const int64_t qb = std::max<int64_t>(256, budget / per_query / 256 * 256);
  return std::min(qb, nq);
}

int run_knn(sd_context* ctx, sd_matrix* X, sd_matrix* Y, const sd_knn_params* p, const KnnPlan& plan,
            int32_t* h_idx, double* h_val, bool keep = false) {
  sd_matrix* ref = plan.selfmode ? X : Y;
  const int method = p->method;
  const int k = plan.k;
  const int64_t nq = X->ncol, nref = ref->ncol;
  HostPrefault pf(h_idx, h_idx ? sizeof(int32_t) * (size_t)nq * (size_t)k : 0, h_val,
                  h_val ? sizeof(double) * (size_t)nq * (size_t)k : 0);
  {
    std::lock_guard<std::mutex> lk(ref->prep_mu);
    SD_TRY(prep_method(ctx, ref, method));
    SD_TRY(sync_stream(ctx));
  }
  if (!plan.selfmode) SD_TRY(prep_method(ctx, X, method));
  SD_TRY(check_cancel(ctx));

  const uint8_t* ref_deg = p->drop_degenerate ? ref->deg.as<uint8_t>() : nullptr;
  const uint8_t* qry_deg_all = p->drop_degenerate ? X->deg.as<uint8_t>() : nullptr;

  DevBuf& out_idx = ctx->scratch[1];
  DevBuf& out_val = ctx->scratch[2];
  SD_TRY(out_idx.ensure(sizeof(int32_t) * (size_t)nq * (size_t)k + 64));
  SD_TRY(out_val.ensure(sizeof(double) * (size_t)nq * (size_t)k + 64));

  int prec = p->precision;
  // cosine re-rank margin: the caller's, else the one tuned on this reference
  int margin = p->margin > 0 ? p->margin : 16;
  if (p->margin <= 0 && prec == SD_PREC_AUTO) {
    std::lock_guard<std::mutex> lk(ref->prep_mu);
    if (ref->tuned_margin > 0) margin = ref->tuned_margin;
  }
  // densified layout: binary -> 0/1 patterns (FP4 nibbles, or FP32 on the
  // plain path); cosine -> unit-normalised values (BF16, or FP32);
  // euclidean and pearson -> raw values (BF16 for non-negative data until a
  // search on this reference falls back too often, else split BF16 hi/lo;
  // or FP32)
  const bool use_f32 = (prec == SD_PREC_FP32);
  bool raw_plain = !ref->any_negative && !X->any_negative;
  if (raw_plain) {
    std::lock_guard<std::mutex> lk(ref->prep_mu);
    raw_plain = !ref->raw_split;
  }
  const int kind = method == SD_BINARY   ? (use_f32 ? kDenseF32Pattern : kDenseF4Pattern)
                   : method == SD_COSINE ? (use_f32 ? kDenseF32Unit : kDenseBf16Unit)
                   : use_f32             ? kDenseF32Raw
                   : raw_plain           ? kDenseBf16Raw
                                         : kDenseBf16Split;
  const int64_t kpad = round_up(std::max<int64_t>(ref->nrow, 1), dense_row_align(kind));
  const bool fits = fused_fits(dense_row_bytes(kind, kpad));
  bool fused_ok_binary = (method == SD_BINARY) && prec != SD_PREC_EXACT &&
                         fused_list_len(k) > 0 && fits;
  bool fused_ok_cosine = (method == SD_COSINE) && prec != SD_PREC_EXACT &&
                         fused_list_len(k + margin) > 0 && nref >= 1 && fits;
  DotBound eb;
  const bool raw_dot = method == SD_EUCLIDEAN || method == SD_PEARSON;
  const bool fused_ok_raw = raw_dot && prec != SD_PREC_EXACT && fused_list_len(k + margin) > 0 &&
                            nref >= 1 && fits && dot_bound(ref, X, kind, &eb);
  bool fused_ok_euclid = fused_ok_raw && method == SD_EUCLIDEAN;
  bool fused_ok_pearson = fused_ok_raw && method == SD_PEARSON;
  if (fused_ok_binary || fused_ok_cosine || fused_ok_euclid || fused_ok_pearson) {
    // SD_PREC_AUTO: the cost model may prefer the merge walk (tiny batches
    // against very sparse references); both paths give identical results
    if (prec == SD_PREC_AUTO &&
        tune_prefer_exact(ctx, kind, nq, nref, kpad,
                          (double)X->nnz / (double)std::max<int64_t>(X->ncol, 1),
                          (double)ref->nnz / (double)std::max<int64_t>(nref, 1),
                          plan.selfmode && nref >= 128 * 128)) {
      fused_ok_binary = false;
      fused_ok_cosine = false;
      fused_ok_euclid = false;
      fused_ok_pearson = false;
    }
  }
  const bool fused_rerank = fused_ok_cosine || fused_ok_euclid || fused_ok_pearson;
  ExactArgs ea = exact_args(ref, X, method, p->dist);
  ea.ref_deg = ref_deg;
  ea.self_mode = plan.selfmode;
  ea.include_self = p->include_self;

  if (fused_ok_binary || fused_rerank) {
    ++ctx->stats.fused_calls;
    ctx->stats.slab_products = kind == kDenseBf16Split ? 3 : 1;
    {
      std::lock_guard<std::mutex> lk(ref->prep_mu);
      SD_TRY(prep_dense(ctx, ref, kind, kpad));
      if (fused_ok_euclid) SD_TRY(prep_sq(ctx, ref));
      SD_TRY(sync_stream(ctx));
    }
    if (!plan.selfmode) {
      SD_TRY(prep_dense(ctx, X, kind, kpad));
      if (fused_ok_euclid) SD_TRY(prep_sq(ctx, X));
    }
    SD_TRY(check_cancel(ctx));

    FusedArgs fa;
    fa.method = method;
    fa.prec = use_f32 ? SD_PREC_FP32 : SD_PREC_BF16;
    fa.kind = kind;
    fa.nref = nref;
    fa.nqry = nq;
    fa.kpad = kpad;
    fa.ref_dense = ref->dense[kind].ptr;
    fa.qry_dense = X->dense[kind].ptr;
    fa.ref_ld = ref->ncol;
    fa.qry_ld = X->ncol;
    fa.ref_p = ref->p.as<int64_t>();
    fa.qry_p = X->p.as<int64_t>();
    fa.ref_deg = ref_deg;
    fa.self_mode = plan.selfmode;
    fa.include_self = p->include_self;
    {
      const double cells = (double)std::max<int64_t>(ref->nrow, 1) * (double)nref;
      const double frac = ref->nnz > 0 ? (double)ref->nnz / cells : 0.0;
      fa.ref_density_bucket = frac > 0 ? (int)std::min(60.0, std::round(-2.0 * std::log2(frac))) : 60;
    }
    if (fused_ok_euclid) {
      fa.ref_aux = ref->sqf.as<float>();
      fa.qry_aux = X->sqf.as<float>();
    } else if (fused_ok_pearson) {
      fa.ref_aux = ref->cmf.as<float>();
      fa.qry_aux = X->cmf.as<float>();
      fa.ref_aux2 = ref->cif.as<float>();
      fa.qry_aux2 = X->cif.as<float>();
    }
    if (fused_ok_euclid || fused_ok_pearson) {
      fa.dot_e0 = eb.e0;
      fa.dot_e1 = eb.e1;
      fa.dot_nonneg = eb.nonneg;
    }
    fa.L = fused_list_len(method == SD_BINARY ? k : k + margin);
    const int L = fa.L;
    // Candidate buffers and lists grow with the query count; past the
    // context's buffer budget the queries run in blocks of columns, each a
    // whole search of its own (a blocked self-search scores every pair from
    // both sides, and a blocked cross search does not share its pass).
    const int64_t qb = fused_query_block(ctx, nq, L);
    if (qb < nq) {
      CrossMirror& xm = ctx->xm;
      if (xm.phase == 1) return SD_OK;  // no shared pass for this cross pair
      if (xm.ready) {
        xm.ready = false;
        xm.cnt.release();
        xm.fkey.release();
        xm.fidx.release();
      }
    }
    ctx->stats.query_blocks += (nq + qb - 1) / qb;
    const FusedArgs fa_all = fa;
    const ExactArgs ea_all = ea;
    const int64_t nq_total = nq;
    for (int64_t q0 = 0; q0 < nq_total; q0 += qb) {
      if (q0 > 0) SD_TRY(check_cancel(ctx));
      // block views: query columns [q0, q0 + nq) of X; results at column q0
      const int64_t nq = std::min(qb, nq_total - q0);
      FusedArgs fa = fa_all;
      ExactArgs ea = ea_all;
      if (q0 > 0 || nq < nq_total) {
        fa.nqry = nq;
        fa.qry_dense = (const char*)fa_all.qry_dense + q0 * kDenseChunk;  // same qry_ld
        fa.qry_p = fa_all.qry_p + q0;
        if (fa_all.qry_aux) fa.qry_aux = fa_all.qry_aux + q0;
        if (fa_all.qry_aux2) fa.qry_aux2 = fa_all.qry_aux2 + q0;
        fa.self_off = q0;
        ea.qry.ncol = nq;
        ea.qry.p = ea_all.qry.p + q0;
        if (ea_all.qry_scale) ea.qry_scale = ea_all.qry_scale + q0;
        if (ea_all.qry_mean) ea.qry_mean = ea_all.qry_mean + q0;
        if (ea_all.qry_css) ea.qry_css = ea_all.qry_css + q0;
        ea.self_off = q0;
      }
      const uint8_t* qry_deg = qry_deg_all ? qry_deg_all + q0 : nullptr;
      int32_t* out_i = out_idx.as<int32_t>() + q0;
      double* out_v = out_val.as<double>() + q0;
      const double* lkey = nullptr;
      const int32_t* lidx = nullptr;
      CandBuffers cb;
      SD_TRY(fused_candidate_lists(ctx, fa, &lkey, &lidx, &cb));
      if (ctx->xm.phase == 1) return SD_OK;  // probe of a cross pair's reverse search
      if (method == SD_BINARY) {
        SD_TRY(finish_exact_lists(ctx, lkey, lidx, nq, L, k, p->dist,
                                  qry_deg, out_i, out_v, nq_total));
      } else {
        // cosine / euclidean / pearson: FP64 re-rank with certificate; queries
        // it cannot clear get a second fused pass over just those columns with
        // the widest list, and only what is still uncertified goes to the exact
        // FP64 slab.
        auto rerank = [&](const RerankArgs& RA) {
          return fused_ok_euclid    ? rerank_euclid(ctx, RA)
                 : fused_ok_pearson ? rerank_pearson(ctx, RA)
                                    : rerank_cosine(ctx, RA);
        };
        DevBuf& fl = ctx->scratch[7];
        DevBuf& fl2 = ctx->scratch[8];
        SD_TRY(fl.ensure(sizeof(int32_t) * (size_t)(nq + 16)));
        int32_t* fail_count = fl.as<int32_t>() + nq;
        SD_TRY(zero_i32(ctx, fail_count, 1));
        RerankArgs R;
        R.ex = ea;
        R.cand_key = lkey;
        R.cand_idx = lidx;
        R.nqry = nq;
        R.L = L;
        R.k = k;
        R.decreasing = plan.dec;
        R.nonneg = !ref->any_negative && !X->any_negative;
        R.in_round = use_f32 ? std::ldexp(1.0, -24) : std::ldexp(1.0, -8) + std::ldexp(1.0, -23);
        R.acc_round = std::ldexp(1.0, -23);
        R.qdeg = qry_deg;
        R.out_idx = out_i;
        R.out_val = out_v;
        R.fail_list = fl.as<int32_t>();
        R.fail_count = fail_count;
        R.qlist = nullptr;
        R.out_ld = nq_total;
        R.map_words = rerank_map_words(ref->nrow);
        SD_TRY(rerank(R));
        if (ctx->capturing) {
          // failures are the caller's to find after replay: an ordinary
          // search answers them
          ctx->cap_fail = fail_count;
          continue;
        }
        int32_t nfail = 0;
        SD_HIP(hipMemcpyAsync(&nfail, fail_count, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        SD_HIP(hipStreamSynchronize(ctx->stream));
        const int32_t* fail_ids = fl.as<int32_t>();
        ctx->stats.last_margin = margin;
        if (prec == SD_PREC_AUTO && p->margin <= 0) {
          // adapt this reference's margin.  After a filtered pass the wider
          // lists come from the candidate buffers and a failed query costs only
          // its second re-rank (64 lanes instead of L): widen when more than 15%
          // fail, narrow by 4 (down to 8) when fewer than 1% do.  Otherwise a
          // failed query costs a second slab pass: widen above 1%, narrow by 8
          // (down to 16) only after a clean search.
          std::lock_guard<std::mutex> lk(ref->prep_mu);
          if (cb.fkey) {
            if ((int64_t)nfail * 100 > 15 * nq)
              ref->tuned_margin = std::min(2 * margin, 48);
            else if ((int64_t)nfail * 100 < nq)
              ref->tuned_margin = std::max(margin - 4, 8);
          } else {
            if ((int64_t)nfail * 100 > nq)
              ref->tuned_margin = std::min(2 * margin, 48);
            else if (nfail == 0)
              ref->tuned_margin = std::max(margin - 8, 16);
          }
        }
        const int Lmax = fused_list_len(64);
        if (nfail > 0 && L < Lmax) {
          SD_TRY(check_cancel(ctx));
          DevBuf& mk2 = ctx->scratch[11];
          DevBuf& mi2 = ctx->scratch[12];
          const double* k2 = nullptr;
          const int32_t* i2 = nullptr;
          if (cb.fkey) {
            // the wider lists straight from the filtered pass's buffers; a
            // failed query whose buffer is not complete for Lmax (fewer
            // entries, or overflowed) gets its list from a list-mode pass over
            // every reference tile instead, placed at its position
            SD_TRY(mk2.ensure(sizeof(double) * (size_t)nfail * Lmax + 64));
            SD_TRY(mi2.ensure(sizeof(int32_t) * (size_t)nfail * Lmax + 64));
            DevBuf& bc = ctx->scratch[15];
            DevBuf& bpos = ctx->scratch[20];
            SD_TRY(bc.ensure(64));
            SD_TRY(bpos.ensure(sizeof(int32_t) * (size_t)(nfail + 16)));
            SD_TRY(zero_i32(ctx, bc.as<int32_t>(), 1));
            SD_TRY(select_lists(ctx, cb.fkey, cb.fidx, cb.cnt, cb.cap, nfail, Lmax, mk2.as<double>(),
                                mi2.as<int32_t>(), bpos.as<int32_t>(), bc.as<int32_t>(), Lmax,
                                fl.as<int32_t>()));
            int32_t nbad = 0;
            SD_HIP(hipMemcpyAsync(&nbad, bc.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
            SD_HIP(hipStreamSynchronize(ctx->stream));
            ctx->stats.escalation_list_queries += nbad;
            if (nbad > 0) {
              DevBuf& bq = ctx->scratch[21];
              SD_TRY(bq.ensure(sizeof(int32_t) * (size_t)(nbad + 16)));
              SD_TRY(gather_i32(ctx, fl.as<int32_t>(), bpos.as<int32_t>(), nbad, bq.as<int32_t>()));
              FusedArgs fb = fa;
              fb.qlist = bq.as<int32_t>();
              fb.nqry = nbad;
              fb.L = Lmax;
              fb.tile_stride = 1;
              const int64_t qt2 = (nbad + 255) / 256;
              fb.splits = (int)std::min<int64_t>(round_up((2 * (int64_t)ctx->cu_count + qt2 - 1) / qt2, 8), 4096);
              const int nl2 = fused_effective_splits(fb);
              DevBuf& ck2 = ctx->scratch[22];
              DevBuf& ci2 = ctx->scratch[23];
              DevBuf& ck3 = ctx->scratch[24];
              DevBuf& ci3 = ctx->scratch[25];
              SD_TRY(ck2.ensure(sizeof(double) * (size_t)nl2 * (size_t)nbad * Lmax + 64));
              SD_TRY(ci2.ensure(sizeof(int32_t) * (size_t)nl2 * (size_t)nbad * Lmax + 64));
              SD_TRY(fused_slab_topl(ctx, fb, ck2.as<double>(), ci2.as<int32_t>()));
              const double* lk = ck2.as<double>();
              const int32_t* li = ci2.as<int32_t>();
              if (nl2 > 1) {
                SD_TRY(ck3.ensure(sizeof(double) * (size_t)nbad * Lmax + 64));
                SD_TRY(ci3.ensure(sizeof(int32_t) * (size_t)nbad * Lmax + 64));
                SD_TRY(merge_lists(ctx, lk, li, nl2, nbad, Lmax, ck3.as<double>(),
                                   ci3.as<int32_t>()));
                lk = ck3.as<double>();
                li = ci3.as<int32_t>();
              }
              SD_TRY(scatter_lists(ctx, lk, li, bpos.as<int32_t>(), nbad, Lmax, mk2.as<double>(),
                                   mi2.as<int32_t>()));
            }
            k2 = mk2.as<double>();
            i2 = mi2.as<int32_t>();
          }
          if (!k2) {
            // second fused pass over just those columns with the widest list
            FusedArgs fb = fa;
            fb.qlist = fl.as<int32_t>();
            fb.nqry = nfail;
            fb.L = Lmax;
            const int64_t qt2 = (nfail + 255) / 256;
            fb.splits = (int)std::min<int64_t>(round_up((2 * (int64_t)ctx->cu_count + qt2 - 1) / qt2, 8), 4096);
            const int nl2 = fused_effective_splits(fb);
            DevBuf& ck2 = ctx->scratch[9];
            DevBuf& ci2 = ctx->scratch[10];
            SD_TRY(ck2.ensure(sizeof(double) * (size_t)nl2 * (size_t)nfail * Lmax + 64));
            SD_TRY(ci2.ensure(sizeof(int32_t) * (size_t)nl2 * (size_t)nfail * Lmax + 64));
            SD_TRY(fused_slab_topl(ctx, fb, ck2.as<double>(), ci2.as<int32_t>()));
            k2 = ck2.as<double>();
            i2 = ci2.as<int32_t>();
            if (nl2 > 1) {
              SD_TRY(mk2.ensure(sizeof(double) * (size_t)nfail * Lmax + 64));
              SD_TRY(mi2.ensure(sizeof(int32_t) * (size_t)nfail * Lmax + 64));
              SD_TRY(merge_lists(ctx, k2, i2, nl2, nfail, Lmax, mk2.as<double>(), mi2.as<int32_t>()));
              k2 = mk2.as<double>();
              i2 = mi2.as<int32_t>();
            }
          } else {
            ++ctx->stats.buffer_escalations;
          }
          SD_TRY(fl2.ensure(sizeof(int32_t) * (size_t)(nfail + 16)));
          int32_t* fail_count2 = fl2.as<int32_t>() + nfail;
          SD_TRY(zero_i32(ctx, fail_count2, 1));
          RerankArgs R2 = R;
          R2.cand_key = k2;
          R2.cand_idx = i2;
          R2.nqry = nfail;
          R2.L = Lmax;
          R2.qlist = fl.as<int32_t>();
          R2.fail_list = fl2.as<int32_t>();
          R2.fail_count = fail_count2;
          SD_TRY(rerank(R2));
          ctx->stats.escalated_queries += nfail;
          SD_HIP(hipMemcpyAsync(&nfail, fail_count2, sizeof(int32_t), hipMemcpyDeviceToHost,
                                ctx->stream));
          SD_HIP(hipStreamSynchronize(ctx->stream));
          fail_ids = fl2.as<int32_t>();
        }
        ctx->stats.fallback_queries += nfail;
        ctx->stats.certified_queries += nq - nfail;
        if (kind == kDenseBf16Raw && (int64_t)nfail * 100 > nq) {
          std::lock_guard<std::mutex> lk(ref->prep_mu);
          ref->raw_split = true;  // later searches on this reference use hi/lo
        }
        if (nfail > 0)
          SD_TRY(exact_knn_columns(ctx, ea, fail_ids, nfail, nq_total, k, plan.dec, 0, out_i,
                                   out_v));
      }
      if (cb.record && g_device_allocs.load(std::memory_order_relaxed) == cb.allocs) {
        // the whole search (pre-pass, filtered pass, selection, re-rank and
        // escalations) per scored pair; a search that had to allocate device
        // memory is not timed (the next one of its class retries)
        SD_HIP(hipStreamSynchronize(ctx->stream));
        const double sec =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - cb.t_start).count();
        sample_record(cb.key, cb.cand, L, sec / cb.pairs);
      }
    }
  } else {
    // exact FP64 path over every query column
    if (ctx->capturing)
      return set_error(SD_EUNSUPPORTED, "internal: the exact path is not captured");
    if (ctx->xm.phase == 1) return SD_OK;  // probe of a cross pair's reverse search
    ++ctx->stats.exact_calls;
    DevBuf& qc = ctx->scratch[7];
    SD_TRY(qc.ensure(sizeof(int32_t) * (size_t)(nq + 16)));
    std::vector<int32_t> iota((size_t)nq);
    for (int64_t j = 0; j < nq; ++j) iota[(size_t)j] = (int32_t)j;
    SD_HIP(hipMemcpyAsync(qc.ptr, iota.data(), sizeof(int32_t) * (size_t)nq, hipMemcpyHostToDevice,
                          ctx->stream));
    SD_TRY(exact_knn_columns(ctx, ea, qc.as<int32_t>(), nq, nq, k, plan.dec, p->block_size,
                             out_idx.as<int32_t>(), out_val.as<double>()));
  }
  SD_TRY(check_cancel(ctx));
  // a degenerate query gets no neighbours (R/sparseKNN.R:298-301)
  if (p->drop_degenerate)
    SD_TRY(fill_degenerate(ctx, out_idx.as<int32_t>(), out_val.as<double>(),
                           X->deg.as<uint8_t>(), nq, k));
  if (h_idx) {
    if (ctx->capturing) return set_error(SD_EUNSUPPORTED, "internal: host results in a capture");
    pf.wait();
    const HostCopy parts[2] = {{h_idx, out_idx.ptr, sizeof(int32_t) * (size_t)nq * (size_t)k},
                               {h_val, out_val.ptr, sizeof(double) * (size_t)nq * (size_t)k}};
    SD_TRY(download(ctx, parts, 2));
  }
  if (keep) {
    // the indices stay on the device for a following SNN call
    ctx->knn_serial = 0;
    SD_TRY(ctx->knn_keep.ensure(sizeof(int32_t) * (size_t)nq * (size_t)k + 64));
    SD_HIP(hipMemcpyAsync(ctx->knn_keep.ptr, out_idx.ptr, sizeof(int32_t) * (size_t)nq * (size_t)k,
                          hipMemcpyDeviceToDevice, ctx->stream));
  }
  SD_TRY(sync_stream(ctx));
  if (keep) {
    ctx->knn_serial = ++ctx->knn_calls;
    ctx->keep_n = nq;
    ctx->keep_k = k;
  }
  return SD_OK;
}

template <class F>
int guarded(F&& f) {
  try {
    clear_error();
    // a HIP error left on this thread by other code would otherwise surface
    // at this call's first launch check
    (void)hipGetLastError();
    return f();
  } catch (const std::bad_alloc&) {
    return set_error(SD_ENOMEM, "host allocation failed");
  } catch (...) {
    return set_error(SD_EDEVICE, "unexpected internal error");
  }
}

// A cancellable call on a context.  A cancel request stays set until the
// call it reaches returns: one sent just before the call starts (the caller's
// interrupt poll can win the race with its worker thread) still stops it, and
// the next call starts clean.
struct CancelScope {
  sd_context* ctx;
  explicit CancelScope(sd_context* c) : ctx(c) {}
  ~CancelScope() {
    if (ctx) ctx->cancel.store(0);
  }
};

}  // namespace

namespace sd {

int knn_device(sd_context* ctx, sd_matrix* X, sd_matrix* Y, const sd_knn_params* p,
               int32_t* k_eff, int32_t* warn, const int32_t** d_idx, const double** d_val) {
  KnnPlan plan;
  int64_t avail = 0;
  SD_TRY(plan_knn(X, Y, p, &plan, &avail));
  *warn = plan.warn;
  *k_eff = plan.k;
  ctx->cap_over = nullptr;
  ctx->cap_fail = nullptr;
  SD_TRY(run_knn(ctx, X, Y, p, plan, nullptr, nullptr));
  *d_idx = ctx->scratch[1].as<int32_t>();
  *d_val = ctx->scratch[2].as<double>();
  return SD_OK;
}

}  // namespace sd

extern "C" {

const char* sd_last_error(void) { return g_err.c_str(); }

const char* sd_version(void) { return "sparsedist_amd 0.1.0 (gfx950)"; }

int sd_context_create(int device, sd_context** out) {
  return guarded([&]() -> int {
    if (!out) return set_error(SD_EINVAL, "out is NULL");
    int ndev = 0;
    SD_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
      return set_error(SD_EINVAL, "device %d not available (%d devices)", device, ndev);
    SD_HIP(hipSetDevice(device));
    auto* c = new sd_context();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_a);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_b);
    if (e != hipSuccess) {
      delete c;
      return set_error(SD_EDEVICE, "context setup failed: %s", hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      c->cu_count = prop.multiProcessorCount;
    const int rc = noalloc_scratch_setup(device, &c->noalloc_scratch);
    if (rc != SD_OK) {
      sd_context_destroy(c);
      return rc;
    }
    *out = c;
    return SD_OK;
  });
}

int sd_context_destroy(sd_context* ctx) {
  if (!ctx) return SD_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  // pooled buffers are freed on the stream, so before it goes
  for (auto& b : ctx->scratch) b.release();
  for (auto& b : ctx->snn_scratch) b.release();
  for (DevBuf* b : {&ctx->xm.thr, &ctx->xm.cnt, &ctx->xm.fkey, &ctx->xm.fidx, &ctx->knn_keep})
    b->release();
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
  if (ctx->ev_a) (void)hipEventDestroy(ctx->ev_a);
  if (ctx->ev_b) (void)hipEventDestroy(ctx->ev_b);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SD_OK;
}

int sd_knn_kept(sd_context* ctx, const int32_t** d_idx, int32_t* n, int32_t* k, int64_t* serial) {
  if (!ctx || !d_idx || !n || !k || !serial) return set_error(SD_EINVAL, "NULL argument");
  const bool kept = ctx->knn_serial > 0;
  *d_idx = kept ? ctx->knn_keep.as<int32_t>() : nullptr;
  *n = kept ? (int32_t)ctx->keep_n : 0;
  *k = kept ? (int32_t)ctx->keep_k : 0;
  *serial = ctx->knn_serial;
  return SD_OK;
}

int sd_cancel(sd_context* ctx) {
  if (!ctx) return SD_EINVAL;
  ctx->cancel.store(1);
  return SD_OK;
}

int sd_context_stats(sd_context* ctx, sd_stats* out, int reset) {
  if (!ctx || !out) return set_error(SD_EINVAL, "NULL argument");
  *out = ctx->stats;
  if (reset) ctx->stats = sd_stats{};
  return SD_OK;
}

int sd_context_set_timing(sd_context* ctx, int on) {
  if (!ctx) return set_error(SD_EINVAL, "NULL context");
  ctx->timing = on != 0;
  return SD_OK;
}

int sd_context_set_buffer_budget(sd_context* ctx, int64_t bytes) {
  return guarded([&]() -> int {
    if (!ctx) return set_error(SD_EINVAL, "NULL context");
    if (bytes < 0) return set_error(SD_EINVAL, "'bytes' must be non-negative.");
    ctx->buffer_budget = bytes;
    return SD_OK;
  });
}

void* sd_context_stream(sd_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int sd_matrix_upload_csc(sd_context* ctx, int64_t nrow, int64_t ncol, const int64_t* colptr,
                         const int32_t* rowind, const double* values, sd_matrix** out) {
  return guarded([&]() -> int {
    if (!ctx || !out) return set_error(SD_EINVAL, "NULL argument");
    SD_HIP(hipSetDevice(ctx->device));
    auto* m = new sd_matrix();
    const int rc = matrix_load(ctx, m, nrow, ncol, colptr, rowind, values);
    if (rc != SD_OK) {
      delete m;
      return rc;
    }
    *out = m;
    return SD_OK;
  });
}

// Freeing waits for work that may still read the buffers: hipFree
// synchronises the device, and pooled buffers are freed in order on their
// stream.  The matrix's own context may be gone already, so it is not used.
int sd_matrix_free(sd_matrix* m) {
  if (!m) return SD_OK;
  (void)hipSetDevice(m->device);
  delete m;
  return SD_OK;
}

int sd_matrix_drop_cache(sd_matrix* m) {
  if (!m) return SD_OK;
  (void)hipSetDevice(m->device);
  m->unit.release();
  m->scale.release();
  m->unit_ssq.release();
  m->rows16.release();
  m->unit_ready = false;
  m->mean.release();
  m->css.release();
  m->cmf.release();
  m->cif.release();
  m->centre_ready = false;
  m->sqn.release();
  m->sqf.release();
  m->sq_ready = false;
  for (int t = 0; t < kDenseKinds; ++t) {
    m->dense[t].release();
    m->dense_kpad[t] = 0;
  }
  return SD_OK;
}

int sd_knn(sd_context* ctx, const sd_matrix* X, const sd_matrix* Y, const sd_knn_params* params,
           int32_t* idx, double* dist, int32_t* k_eff, int32_t* warn) {
  return guarded([&]() -> int {
    if (!ctx || !X || !params || !idx || !dist || !k_eff || !warn)
      return set_error(SD_EINVAL, "NULL argument");
    SD_HIP(hipSetDevice(ctx->device));
    CancelScope scope(ctx);
    KnnPlan plan;
    int64_t avail = 0;
    SD_TRY(plan_knn(X, Y, params, &plan, &avail));
    *warn = plan.warn;
    *k_eff = plan.k;
    return run_knn(ctx, const_cast<sd_matrix*>(X), const_cast<sd_matrix*>(Y), params, plan, idx,
                   dist, plan.selfmode);
  });
}

int sd_knn_cross(sd_context* ctx, const sd_matrix* X, const sd_matrix* Y,
                 const sd_knn_params* params, int32_t* fwd_idx, double* fwd_dist, int32_t* k_fwd,
                 int32_t* rev_idx, double* rev_dist, int32_t* k_rev, int32_t* warn) {
  return guarded([&]() -> int {
    if (!ctx || !X || !Y || !params || !fwd_idx || !fwd_dist || !rev_idx || !rev_dist || !k_fwd ||
        !k_rev || !warn)
      return set_error(SD_EINVAL, "NULL argument");
    SD_HIP(hipSetDevice(ctx->device));
    CancelScope scope(ctx);
    if (params->k < 1)
      return set_error(SD_EINVAL,
                       "'k' must be a positive integer no larger than .Machine$integer.max.");
    int dec = 0;
    SD_TRY(method_direction(params->method, params->dist, &dec));
    if (X->nrow != Y->nrow)
      return set_error(SD_EINVAL, "'X' and 'Y' must have the same number of rows.");
    if (X->ncol < 1) return set_error(SD_EINVAL, "'X' must have at least one column.");
    if (Y->ncol < 1) return set_error(SD_EINVAL, "'Y' must have at least one column.");
    // per-direction clamp (R/sparseKNNCross.R:117-127)
    KnnPlan pf, pr;
    int64_t av_f = 0, av_r = 0;
    *warn = 0;
    int rc = plan_knn(X, Y, params, &pf, &av_f);
    if (rc == SD_ENOCAND)
      return set_error(SD_ENOCAND, "No candidate neighbours available for the forward direction.");
    SD_TRY(rc);
    rc = plan_knn(Y, X, params, &pr, &av_r);
    if (rc == SD_ENOCAND)
      return set_error(SD_ENOCAND, "No candidate neighbours available for the reverse direction.");
    SD_TRY(rc);
    *warn = (pf.warn ? 1 : 0) | (pr.warn ? 2 : 0);
    *k_fwd = pf.k;
    *k_rev = pr.k;
    sd_matrix* Xm = const_cast<sd_matrix*>(X);
    sd_matrix* Ym = const_cast<sd_matrix*>(Y);
    // One filtered pass for both directions when both are large enough for
    // one (at least 128 reference tiles of 128 columns each way): probe the
    // reverse search (its thresholds and buffers), run the forward search,
    // whose filtered pass also feeds the reverse buffers, and resume the
    // reverse search from them.  When the searches take other paths (or
    // other layouts), each runs on its own as before.
    CrossMirror& xm = ctx->xm;
    struct Reset {
      CrossMirror& m;
      ~Reset() {
        m.phase = 0;
        m.ready = false;
        m.target.filled = false;
      }
    } reset{xm};
    xm.ready = false;
    xm.target.filled = false;
    if (params->precision != SD_PREC_EXACT && X->ncol >= 128 * 128 && Y->ncol >= 128 * 128) {
      const int64_t fused = ctx->stats.fused_calls, blocks = ctx->stats.query_blocks;
      xm.phase = 1;
      SD_TRY(run_knn(ctx, Ym, Xm, params, pr, rev_idx, rev_dist));  // writes no results
      ctx->stats.fused_calls = fused;  // counted when it resumes
      ctx->stats.query_blocks = blocks;
      xm.phase = 2;
    }
    // the forward results wait in host memory until the reverse search has
    // succeeded, so a cancelled or failed call leaves every output untouched
    std::vector<int32_t> fi((size_t)X->ncol * (size_t)pf.k);
    std::vector<double> fd((size_t)X->ncol * (size_t)pf.k);
    SD_TRY(run_knn(ctx, Xm, Ym, params, pf, fi.data(), fd.data()));
    if (xm.phase == 2) xm.phase = 3;
    SD_TRY(run_knn(ctx, Ym, Xm, params, pr, rev_idx, rev_dist));
    std::copy(fi.begin(), fi.end(), fwd_idx);
    std::copy(fd.begin(), fd.end(), fwd_dist);
    return SD_OK;
  });
}

int sd_slab(sd_context* ctx, int32_t method, const sd_matrix* X, const sd_matrix* Y, int32_t dist,
            double* D) {
  return guarded([&]() -> int {
    if (!ctx || !X || !Y || !D) return set_error(SD_EINVAL, "NULL argument");
    SD_HIP(hipSetDevice(ctx->device));
    if (X->nrow != Y->nrow)
      return set_error(SD_EINVAL, "Mismatched row dimensions of m (%lld) and m2 (%lld).",
                       (long long)X->nrow, (long long)Y->nrow);
    int dec;
    if (method == SD_EUCLIDEAN || method == SD_MANHATTAN || method == SD_JS) dist = 1;
    SD_TRY(method_direction(method, dist, &dec));
    sd_matrix* m = const_cast<sd_matrix*>(X);
    sd_matrix* m2 = const_cast<sd_matrix*>(Y);
    SD_TRY(prep_method(ctx, m, method));
    SD_TRY(prep_method(ctx, m2, method));
    ExactArgs a = exact_args(m, m2, method, dist);
    const int64_t nr = m->ncol, nc = m2->ncol;
    if (nr == 0 || nc == 0) return SD_OK;
    DevBuf dD, qc;
    const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(nc, ((int64_t)1 << 29) / std::max<int64_t>(nr, 1)));
    SD_TRY(dD.alloc(sizeof(double) * (size_t)nr * (size_t)nb));
    SD_TRY(qc.alloc(sizeof(int32_t) * (size_t)nc));
    std::vector<int32_t> iota((size_t)nc);
    for (int64_t j = 0; j < nc; ++j) iota[(size_t)j] = (int32_t)j;
    SD_TRY(copy_wait(qc.ptr, iota.data(), sizeof(int32_t) * (size_t)nc,
                     hipMemcpyHostToDevice, ctx->stream));
    for (int64_t c0 = 0; c0 < nc; c0 += nb) {
      const int64_t n = std::min(nb, nc - c0);
      SD_TRY(exact_slab(ctx, a, qc.as<int32_t>() + c0, n, dD.as<double>(), nr, 0));
      SD_HIP(hipMemcpyAsync(D + c0 * nr, dD.ptr, sizeof(double) * (size_t)nr * (size_t)n,
                            hipMemcpyDeviceToHost, ctx->stream));
      SD_HIP(hipStreamSynchronize(ctx->stream));
    }
    return SD_OK;
  });
}

int sd_topk_block(sd_context* ctx, const double* D, int64_t nrow, int64_t ncol, int32_t k,
                  int32_t decreasing, const int32_t* self_row, int32_t* idx, double* val) {
  return guarded([&]() -> int {
    if (!ctx || !D || !self_row || !idx || !val) return set_error(SD_EINVAL, "NULL argument");
    SD_HIP(hipSetDevice(ctx->device));
    if (k < 1) return set_error(SD_EINVAL, "k must be >= 1.");
    std::vector<int32_t> skip((size_t)ncol);
    for (int64_t b = 0; b < ncol; ++b) {
      const int32_t sr = self_row[b];
      if (sr == kNaInt || sr < 0 || (int64_t)sr > nrow)
        return set_error(SD_EINVAL, "self_row entries must be 0 or valid 1-based rows of D.");
      skip[(size_t)b] = sr - 1;
    }
    if (ncol == 0) return SD_OK;
    DevBuf dD, ds, di, dv;
    SD_TRY(dD.alloc(sizeof(double) * (size_t)nrow * (size_t)ncol));
    SD_TRY(ds.alloc(sizeof(int32_t) * (size_t)ncol));
    SD_TRY(di.alloc(sizeof(int32_t) * (size_t)ncol * (size_t)k));
    SD_TRY(dv.alloc(sizeof(double) * (size_t)ncol * (size_t)k));
    SD_TRY(copy_wait(dD.ptr, D, sizeof(double) * (size_t)nrow * (size_t)ncol,
                     hipMemcpyHostToDevice, ctx->stream));
    SD_TRY(copy_wait(ds.ptr, skip.data(), sizeof(int32_t) * (size_t)ncol,
                     hipMemcpyHostToDevice, ctx->stream));
    SD_TRY(topk_dense(ctx, dD.as<double>(), nrow, ncol, nrow, k, decreasing, ds.as<int32_t>(),
                      di.as<int32_t>(), dv.as<double>(), ncol, nullptr));
    SD_HIP(hipStreamSynchronize(ctx->stream));
    SD_TRY(copy_wait(idx, di.ptr, sizeof(int32_t) * (size_t)ncol * (size_t)k,
                     hipMemcpyDeviceToHost, ctx->stream));
    SD_TRY(copy_wait(val, dv.ptr, sizeof(double) * (size_t)ncol * (size_t)k,
                     hipMemcpyDeviceToHost, ctx->stream));
    return SD_OK;
  });
}

int sd_topk_rows_accum(sd_context* ctx, const double* D, int64_t nrow, int64_t ncol, int32_t k,
                       int32_t decreasing, int32_t* best_idx, double* best_val,
                       int64_t col_offset, const int32_t* self_col) {
  return guarded([&]() -> int {
    if (!ctx || !D || !best_idx || !best_val || !self_col)
      return set_error(SD_EINVAL, "NULL argument");
    SD_HIP(hipSetDevice(ctx->device));
    if (k < 1) return set_error(SD_EINVAL, "k must be >= 1.");
    if (col_offset < 0) return set_error(SD_EINVAL, "col_offset must be >= 0.");
    for (int64_t i = 0; i < nrow; ++i)
      if (self_col[i] == kNaInt || self_col[i] < 0)
        return set_error(SD_EINVAL,
                         "self_col entries must be 0 or a positive 1-based query index.");
    if (nrow == 0) return SD_OK;
    DevBuf dD, dsc, di, dv;
    SD_TRY(dD.alloc(sizeof(double) * (size_t)nrow * (size_t)std::max<int64_t>(ncol, 1)));
    SD_TRY(dsc.alloc(sizeof(int32_t) * (size_t)nrow));
    SD_TRY(di.alloc(sizeof(int32_t) * (size_t)nrow * (size_t)k));
    SD_TRY(dv.alloc(sizeof(double) * (size_t)nrow * (size_t)k));
    if (ncol > 0)
      SD_TRY(copy_wait(dD.ptr, D, sizeof(double) * (size_t)nrow * (size_t)ncol,
                       hipMemcpyHostToDevice, ctx->stream));
    SD_TRY(copy_wait(dsc.ptr, self_col, sizeof(int32_t) * (size_t)nrow,
                     hipMemcpyHostToDevice, ctx->stream));
    SD_TRY(copy_wait(di.ptr, best_idx, sizeof(int32_t) * (size_t)nrow * (size_t)k,
                     hipMemcpyHostToDevice, ctx->stream));
    SD_TRY(copy_wait(dv.ptr, best_val, sizeof(double) * (size_t)nrow * (size_t)k,
                     hipMemcpyHostToDevice, ctx->stream));
    SD_TRY(topk_rows(ctx, dD.as<double>(), nrow, ncol, k, decreasing, di.as<int32_t>(),
                     dv.as<double>(), col_offset, dsc.as<int32_t>()));
    SD_HIP(hipStreamSynchronize(ctx->stream));
    SD_TRY(copy_wait(best_idx, di.ptr, sizeof(int32_t) * (size_t)nrow * (size_t)k,
                     hipMemcpyDeviceToHost, ctx->stream));
    SD_TRY(copy_wait(best_val, dv.ptr, sizeof(double) * (size_t)nrow * (size_t)k,
                     hipMemcpyDeviceToHost, ctx->stream));
    return SD_OK;
  });
}

int sd_snn_jaccard(sd_context* ctx, const int32_t* idx, int32_t n, int32_t k,
                   int32_t include_self, double* w) {
  return guarded([&]() -> int {
    if (!ctx || !idx || !w) return set_error(SD_EINVAL, "NULL argument");
    if (n < 1 || k < 1) return set_error(SD_EINVAL, "'idx' must have at least one row and column.");
    SD_HIP(hipSetDevice(ctx->device));
    DevBuf di, dw;
    const size_t total = (size_t)n * (size_t)k;
    SD_TRY(di.alloc(sizeof(int32_t) * total));
    SD_TRY(dw.alloc(sizeof(double) * total));
    SD_TRY(copy_wait(di.ptr, idx, sizeof(int32_t) * total, hipMemcpyHostToDevice, ctx->stream));
    SD_TRY(snn_build(ctx, di.as<int32_t>(), n, k, include_self, 0.0, 0, 0, nullptr, dw.as<double>()));
    SD_TRY(copy_wait(w, dw.ptr, sizeof(double) * total, hipMemcpyDeviceToHost, ctx->stream));
    return SD_OK;
  });
}

int sd_snn(sd_context* ctx, const int32_t* idx, int32_t n, int32_t k, int32_t include_self,
           double prune, int32_t symmetrise, int32_t edges, int32_t idx_on_device,
           sd_graph** out) {
  return guarded([&]() -> int {
    if (!ctx || !idx || !out) return set_error(SD_EINVAL, "NULL argument");
    if (n < 1 || k < 1) return set_error(SD_EINVAL, "'idx' must have at least one row and column.");
    if (!(prune >= 0.0 && prune <= 1.0))
      return set_error(SD_EINVAL, "'prune' must be a finite number between 0 and 1.");
    if (edges != 0 && edges != 1) return set_error(SD_EINVAL, "edges must be 0 (knn) or 1 (shared)");
    SD_HIP(hipSetDevice(ctx->device));
    CancelScope scope(ctx);
    DevBuf& di = ctx->snn_scratch[0];
    const int32_t* d_idx = idx;
    if (idx_on_device) {
      ++ctx->stats.snn_device_inputs;
    } else {
      const size_t total = (size_t)n * (size_t)k;
      SD_TRY(di.ensure(sizeof(int32_t) * total));
      SD_HIP(hipMemcpyAsync(di.ptr, idx, sizeof(int32_t) * total, hipMemcpyHostToDevice,
                            ctx->stream));
      d_idx = di.as<int32_t>();
    }
    auto* g = new sd_graph();
    g->device = ctx->device;
    int rc = snn_build(ctx, d_idx, n, k, include_self, prune, symmetrise, edges, g, nullptr);
    if (rc != SD_OK) {
      delete g;
      return rc;
    }
    *out = g;
    return SD_OK;
  });
}

int sd_graph_info(const sd_graph* g, int64_t* n, int64_t* nnz) {
  if (!g || !n || !nnz) return set_error(SD_EINVAL, "NULL argument");
  *n = g->n;
  *nnz = g->nnz;
  return SD_OK;
}

int sd_graph_copy(const sd_graph* g, int64_t* colptr, int32_t* rowind, double* val) {
  return guarded([&]() -> int {
    if (!g || !colptr) return set_error(SD_EINVAL, "NULL argument");
    if (g->nnz > 0 && (!rowind || !val)) return set_error(SD_EINVAL, "NULL argument");
    hipStream_t s = nullptr;
    SD_TRY(transfer_stream(g->device, &s));
    // the three arrays in one round trip; the stream is drained before any
    // return, so no copy is left writing the caller's arrays
    hipError_t e = hipMemcpyAsync(colptr, g->colptr.ptr, sizeof(int64_t) * (size_t)(g->n + 1),
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && g->nnz > 0)
      e = hipMemcpyAsync(rowind, g->rowind.ptr, sizeof(int32_t) * (size_t)g->nnz,
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && g->nnz > 0)
      e = hipMemcpyAsync(val, g->val.ptr, sizeof(double) * (size_t)g->nnz,
                         hipMemcpyDeviceToHost, s);
    const hipError_t es = hipStreamSynchronize(s);
    SD_HIP(e);
    SD_HIP(es);
    return SD_OK;
  });
}

int sd_graph_free(sd_graph* g) {
  delete g;
  return SD_OK;
}

int sd_detect_cores(void) {
  const unsigned n = std::thread::hardware_concurrency();
  return (n < 2) ? 1 : (int)n - 1;
}

}  // extern "C"
