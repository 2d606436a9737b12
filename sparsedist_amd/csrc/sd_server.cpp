// sd_server.cpp -- request loop over a resident reference matrix.
//
// Requests are sd_knn(X_request, Y) calls (plus, on request, the
// shared-neighbour graph of the query columns) served by a fixed set of
// worker slots.  Each slot owns an sd_context, i.e. its own HIP stream and
// scratch memory, so the work of different requests overlaps on the device.
// A request is cut into work items of at most chunk_cols query columns;
// a free slot always takes the next item of the admitted request with the
// fewest unscheduled columns, so small requests overtake large ones.
// Per-query results are independent of the chunking and of the requests
// sharing the device, so each result equals the one-shot sd_knn call.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <tuple>

#include "sd_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

template <class F>
int guarded_call(F&& f) {
  try {
    sd::clear_error();
    return f();
  } catch (const std::bad_alloc&) {
    return sd::set_error(SD_ENOMEM, "host allocation failed");
  } catch (...) {
    return sd::set_error(SD_EDEVICE, "unexpected internal error");
  }
}

struct Request {
  int64_t ticket = 0;
  // inputs: host copies, or the caller's arrays when borrowed
  int64_t nrow = 0, ncol = 0, nnz = 0;
  std::vector<int64_t> own_p;
  std::vector<int32_t> own_i;
  std::vector<double> own_x;
  const int64_t* colptr = nullptr;
  const int32_t* rowind = nullptr;
  const double* values = nullptr;
  sd_knn_params knn{};
  int snn = 0;
  double prune = 0.0;
  int64_t reserve = 0;  // bytes reserved while admitted
  bool in_lane = false;  // admitted into the small-request lane
  int bypassed = 0;     // requests admitted ahead of it while it waited
  // progress (guarded by the server mutex)
  int64_t next_col = 0;   // first unscheduled column
  int64_t done_cols = 0;  // columns whose results are in place
  int inflight = 0;       // items taken and not yet returned
  bool admitted = false, finished = false;
  int64_t finish_order = 0;
  int status = SD_OK;
  std::string error;
  int32_t k_eff = 0, warn = 0;
  // results: ncol x k column-major, compacted to k_eff columns at the end
  std::vector<int32_t> idx;
  std::vector<double> dist;
  sd_graph* graph = nullptr;
  Clock::time_point t_submit, t_admit, t_done;
};

// ---- graph items -------------------------------------------------------------
// A work item of at most kGraphCols query columns (a small request, or small
// requests packed together) runs as one replay of a HIP graph.  The slot
// pads the item to exactly kGraphCols columns with copies of one of its
// non-empty columns (every query's result is independent of the others, so
// padding changes no answer), loads it into query buffers sized once for
// the widest item, and replays the launch sequence captured for the
// search's parameters: query preparation, threshold pre-pass, filtered pass,
// selection, FP64 re-rank with the widest list and certificate, results.
// The filtered pass's overflow count and the certificate's failure count
// come back with the results; when either is non-zero the item runs as an
// ordinary search instead.  A graph is captured the first time its
// parameters occur, right after an ordinary run of the same item has grown
// every buffer it needs (that run answers the item), and again whenever a
// buffer it reads has moved.
constexpr int64_t kGraphCols = 256;

struct GraphKey {
  int32_t k, method, dist, drop_degenerate, precision, margin;
  bool operator<(const GraphKey& o) const {
    return std::tie(k, method, dist, drop_degenerate, precision, margin) <
           std::tie(o.k, o.method, o.dist, o.drop_degenerate, o.precision, o.margin);
  }
};

struct GraphEntry {
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> ptrs;  // device buffers the launches were captured with
  const int32_t* over = nullptr;  // counters the replay leaves (null: none)
  const int32_t* fail = nullptr;
  const int32_t* out_idx = nullptr;
  const double* out_val = nullptr;
  int32_t k_eff = 0, warn = 0;
  bool broken = false;  // the search cannot be captured: ordinary searches only
};

struct GraphSlot {
  sd_context* ctx = nullptr;  // its own context: no other work grows its buffers
  sd_matrix* X = nullptr;     // the padded query columns
  std::map<GraphKey, GraphEntry> graphs;
};

}  // namespace

struct sd_server {
  int device = 0;
  const sd_matrix* Y = nullptr;
  int slots = 4;
  int64_t chunk = 4096;
  int64_t budget = 0;
  int64_t max_pending = 0;
  std::vector<sd_context*> ctx;
  // per slot: the query matrix of its current item, reloaded in place so that
  // items allocate no device memory once its buffers have grown
  std::vector<sd_matrix*> qmat;
  // per slot: captured searches of small items (empty with no_graphs)
  std::vector<GraphSlot> gslot;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_gone;
  bool stop = false;
  bool paused = false;
  int64_t finished_count = 0;
  int waiters = 0;  // callers blocked in sd_server_wait (destroy waits for them)
  int64_t next_ticket = 1;
  std::map<int64_t, std::unique_ptr<Request>> all;
  std::deque<Request*> pending;
  std::vector<Request*> active;
  int64_t reserved = 0;
  int64_t reserved_lane = 0;  // reservations of requests admitted into the small lane
  sd_server_stats st{};
  std::string last_msg;  // storage for sd_server_error's return value
};

namespace {

// Device bytes one request may hold while admitted: one work item's
// densified query columns, candidate buffers and staging (the slot scratch
// grows to this), its results, and the SNN workspace (sets, inverted index
// and up to n k^2 candidate pairs, sorted with keys and values).
int64_t reserve_estimate(const sd_server* s, const Request& r) {
  const int64_t cols = std::min<int64_t>(r.ncol, s->chunk);
  const int64_t kpad = (std::max<int64_t>(r.nrow, 1) + 511) / 512 * 512;  // widest alignment
  const int64_t k = std::max(r.knn.k, 1);
  const int64_t L = 64;
  const int64_t cap = 4096;
  const int64_t chunk_nnz = r.ncol > 0 ? (r.nnz + r.ncol - 1) / r.ncol * cols : 0;
  int64_t b = cols * kpad * 2              // densified query columns (BF16)
              + cols * cap * 12            // candidate buffer (key + index)
              + cols * L * 12 * 4          // candidate lists and merges
              + chunk_nnz * 28 + cols * 64  // CSC copy, unit values, flags
              + cols * k * 12;             // results of one item
  if (r.snn) b += r.ncol * k * (k + 1) * 4 + r.ncol * k * k * 40 + r.ncol * 64;
  return b + (1 << 20);
}

// A waiting request that does not fit lets later requests that fit pass it
// until it has waited kAgeMs or been passed kMaxBypass times; from then on
// nothing behind it is admitted first, so no request waits for ever.
constexpr double kAgeMs = 2000.0;
constexpr int kMaxBypass = 64;

// Small requests (at most kLaneCols columns) that the budget cannot take
// yet, or that wait behind a request which has stopped the queue, may use a
// lane of their own: up to budget / kLaneShare bytes beyond the budget.  A
// large request that holds the queue then delays only other large ones; the
// lane is bounded, so it delays no large request either.
constexpr int64_t kLaneCols = 1000;
constexpr int64_t kLaneShare = 16;

// Admit waiting requests in arrival order while their reservations fit,
// letting fitting requests pass blocked ones that have not aged yet; then
// small requests still waiting into the lane while it has room.
void admit_locked(sd_server* s) {
  const auto now = Clock::now();
  auto admit = [&](std::deque<Request*>::iterator it, bool lane) {
    Request* r = *it;
    // every request still waiting ahead of r is passed once more
    for (auto jt = s->pending.begin(); jt != it; ++jt) ++(*jt)->bypassed;
    if (lane) {
      s->reserved_lane += r->reserve;
      ++s->st.lane_admitted;
    } else {
      s->reserved += r->reserve;
    }
    r->in_lane = lane;
    r->admitted = true;
    r->t_admit = now;
    s->active.push_back(r);
    ++s->st.admitted;
    return s->pending.erase(it);
  };
  for (auto it = s->pending.begin(); it != s->pending.end();) {
    Request* r = *it;
    if (s->reserved + r->reserve <= s->budget) {
      it = admit(it, false);
      continue;
    }
    if (r->bypassed >= kMaxBypass || ms_between(r->t_submit, now) >= kAgeMs) break;
    ++it;
  }
  for (auto it = s->pending.begin(); it != s->pending.end();) {
    Request* r = *it;
    if (r->ncol <= kLaneCols && s->reserved_lane + r->reserve <= s->budget / kLaneShare)
      it = admit(it, true);
    else
      ++it;
  }
}

void finish_locked(sd_server* s, Request* r) {
  r->finished = true;
  r->finish_order = ++s->finished_count;
  r->t_done = Clock::now();
  (r->in_lane ? s->reserved_lane : s->reserved) -= r->reserve;
  s->active.erase(std::remove(s->active.begin(), s->active.end(), r), s->active.end());
  ++s->st.completed;
  s->st.query_cols += r->ncol;
  admit_locked(s);
  s->cv_done.notify_all();
  s->cv_work.notify_all();
}

// The admitted request with the fewest unscheduled columns (ties: earliest).
Request* pick_locked(sd_server* s) {
  Request* best = nullptr;
  for (Request* r : s->active) {
    if (r->next_col >= r->ncol) continue;
    const int64_t left = r->ncol - r->next_col;
    if (!best || left < best->ncol - best->next_col ||
        (left == best->ncol - best->next_col && r->ticket < best->ticket))
      best = r;
  }
  return best;
}

// A piece of a work item: query columns [c0, c1) of request r.
struct Seg {
  Request* r;
  int64_t c0, c1;
};

// Requests whose columns may share one search: every parameter that can
// change a result is equal (the block size is a memory hint only).
bool same_search(const sd_knn_params& a, const sd_knn_params& b) {
  return a.k == b.k && a.method == b.method && a.dist == b.dist &&
         a.drop_degenerate == b.drop_degenerate && a.precision == b.precision &&
         a.margin == b.margin;
}

// Results of one search over the pieces' concatenated columns (column-major
// with leading dimension ld, k_eff columns) go to each piece's request.
void scatter_results(sd_server* s, const std::vector<Seg>& segs, const int32_t* idx,
                     const double* dist, int64_t ld, int32_t keff, int32_t warn) {
  // columns [c0, c1) of each request belong to this item alone, so no lock
  // is needed for the copies; k_eff is a property of (Y, k) and equal for
  // every item.
  int64_t off = 0;
  for (const Seg& g : segs) {
    const int64_t n = g.c1 - g.c0;
    for (int32_t t = 0; t < keff; ++t) {
      std::copy(idx + (size_t)t * ld + off, idx + (size_t)t * ld + off + n,
                g.r->idx.begin() + (size_t)t * g.r->ncol + g.c0);
      std::copy(dist + (size_t)t * ld + off, dist + (size_t)t * ld + off + n,
                g.r->dist.begin() + (size_t)t * g.r->ncol + g.c0);
    }
    off += n;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  for (const Seg& g : segs) {
    g.r->k_eff = keff;
    g.r->warn = warn;
  }
}

bool graph_eligible(const sd_server* s, const sd_knn_params& k, int64_t ncol) {
  return !s->gslot.empty() && ncol <= kGraphCols &&
         (k.method == SD_COSINE || k.method == SD_BINARY) &&
         (k.precision == SD_PREC_AUTO || k.precision == SD_PREC_BF16);
}

// Device buffers a captured search reads or writes: the graph context's
// scratch and the query and reference matrices' arrays and layouts.
std::vector<const void*> graph_ptrs(const sd_context* c, const sd_matrix* X, const sd_matrix* Y) {
  std::vector<const void*> v;
  for (const sd::DevBuf& b : c->scratch) v.push_back(b.ptr);
  for (const sd_matrix* m : {X, Y}) {
    for (const sd::DevBuf* b : {&m->p, &m->i, &m->x, &m->deg, &m->unit, &m->scale, &m->unit_ssq,
                                &m->mean, &m->css, &m->cmf, &m->cif, &m->sqn, &m->sqf})
      v.push_back(b->ptr);
    for (const sd::DevBuf& b : m->dense) v.push_back(b.ptr);
  }
  // the uint16 row indices are read on the reference side only (the cosine
  // re-rank); the query's copy moves as it grows and is not watched
  v.push_back((Y ? Y : X)->rows16.ptr);
  return v;
}

// The item from its graph (see GraphSlot).  *done = false leaves the item to
// an ordinary search: only empty columns, counters that need a second pass,
// or any failure here (an ordinary search then reports it to the right
// request).
int run_graph_item(sd_server* s, int slot, const std::vector<Seg>& segs, bool* done) {
  *done = false;
  GraphSlot& gs = s->gslot[(size_t)slot];
  sd_context* ctx = gs.ctx;
  const Request* r0 = segs[0].r;
  // the item's columns, then copies of its first column with a non-zero value
  std::vector<int64_t> p(1, 0);
  std::vector<int32_t> ri;
  std::vector<double> vx;
  const Request* pad_r = nullptr;
  int64_t pad_c = -1;
  for (const Seg& g : segs)
    for (int64_t c = g.c0; c < g.c1; ++c) {
      const int64_t b0 = g.r->colptr[c], b1 = g.r->colptr[c + 1];
      ri.insert(ri.end(), g.r->rowind + b0, g.r->rowind + b1);
      vx.insert(vx.end(), g.r->values + b0, g.r->values + b1);
      p.push_back((int64_t)ri.size());
      for (int64_t e = b0; pad_c < 0 && e < b1; ++e)
        if (g.r->values[e] != 0.0) {
          pad_r = g.r;
          pad_c = c;
        }
    }
  if (pad_c < 0) return SD_OK;
  const int64_t nc = (int64_t)p.size() - 1;
  {
    const int64_t b0 = pad_r->colptr[pad_c], b1 = pad_r->colptr[pad_c + 1];
    while ((int64_t)p.size() - 1 < kGraphCols) {
      ri.insert(ri.end(), pad_r->rowind + b0, pad_r->rowind + b1);
      vx.insert(vx.end(), pad_r->values + b0, pad_r->values + b1);
      p.push_back((int64_t)ri.size());
    }
  }
  sd_matrix* X = gs.X;
  sd_matrix* Y = const_cast<sd_matrix*>(s->Y);
  // query buffers sized once for the widest item, so that loading an item
  // never moves them
  const size_t cap_nnz = (size_t)std::max<int64_t>(Y->nrow, 1) * kGraphCols + 1;
  SD_TRY(X->i.ensure(sizeof(int32_t) * cap_nnz));
  SD_TRY(X->x.ensure(sizeof(double) * cap_nnz));
  SD_TRY(X->unit.ensure(sizeof(double) * cap_nnz));
  SD_TRY(sd::matrix_load(ctx, X, r0->nrow, kGraphCols, p.data(), ri.data(), vx.data()));
  sd_knn_params prm = r0->knn;
  prm.include_self = 0;
  // cosine: the widest candidate list, so that the certificate clears nearly
  // every query in the first re-rank (results do not depend on the margin)
  if (prm.method == SD_COSINE && prm.margin <= 0 && prm.k <= 56) prm.margin = 64 - prm.k;
  const GraphKey key{prm.k, prm.method, prm.dist, prm.drop_degenerate, prm.precision, prm.margin};
  GraphEntry& e = gs.graphs[key];
  if (e.broken) return SD_OK;
  std::vector<int32_t> idx;
  std::vector<double> dist;
  auto fetch = [&](const int32_t* di, const double* dv, int32_t keff) -> int {
    idx.resize((size_t)kGraphCols * (size_t)keff);
    dist.resize((size_t)kGraphCols * (size_t)keff);
    SD_HIP(hipMemcpyAsync(idx.data(), di, sizeof(int32_t) * idx.size(), hipMemcpyDeviceToHost,
                          ctx->stream));
    SD_HIP(hipMemcpyAsync(dist.data(), dv, sizeof(double) * dist.size(), hipMemcpyDeviceToHost,
                          ctx->stream));
    return SD_OK;
  };
  if (e.exec && e.ptrs == graph_ptrs(ctx, X, Y)) {
    SD_HIP(hipGraphLaunch(e.exec, ctx->stream));
    int32_t cnt[2] = {0, 0};
    SD_TRY(fetch(e.out_idx, e.out_val, e.k_eff));
    if (e.over)
      SD_HIP(hipMemcpyAsync(&cnt[0], e.over, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    if (e.fail)
      SD_HIP(hipMemcpyAsync(&cnt[1], e.fail, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    SD_HIP(hipStreamSynchronize(ctx->stream));
    if (cnt[0] != 0 || cnt[1] != 0) {
      std::lock_guard<std::mutex> lk(s->mu);
      ++s->st.graph_fallbacks;
      return SD_OK;
    }
    scatter_results(s, segs, idx.data(), dist.data(), kGraphCols, e.k_eff, e.warn);
    std::lock_guard<std::mutex> lk(s->mu);
    ++s->st.graph_items;
    *done = true;
    return SD_OK;
  }
  // An ordinary run answers the item and grows every buffer the search
  // needs; then the same search is captured (capture runs nothing).
  if (e.exec) {
    (void)hipGraphExecDestroy(e.exec);
    e.exec = nullptr;
  }
  int32_t keff = 0, warn = 0;
  const int32_t* di = nullptr;
  const double* dv = nullptr;
  SD_TRY(sd::knn_device(ctx, X, Y, &prm, &keff, &warn, &di, &dv));
  SD_TRY(fetch(di, dv, keff));
  SD_HIP(hipStreamSynchronize(ctx->stream));
  scatter_results(s, segs, idx.data(), dist.data(), kGraphCols, keff, warn);
  *done = true;
  const std::vector<const void*> before = graph_ptrs(ctx, X, Y);
  sd::matrix_forget_derived(X);  // the capture records the query preparation too
  hipGraph_t g = nullptr;
  int rc = SD_EDEVICE;
  if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    ctx->capturing = true;
    try {
      rc = sd::knn_device(ctx, X, Y, &prm, &keff, &warn, &di, &dv);
    } catch (...) {
      rc = SD_EDEVICE;
    }
    ctx->capturing = false;
    if (hipStreamEndCapture(ctx->stream, &g) != hipSuccess && rc == SD_OK) rc = SD_EDEVICE;
  }
  if (rc == SD_OK && g && graph_ptrs(ctx, X, Y) == before &&
      hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0) == hipSuccess) {
    e.ptrs = before;
    e.over = ctx->cap_over;
    e.fail = ctx->cap_fail;
    e.out_idx = di;
    e.out_val = dv;
    e.k_eff = keff;
    e.warn = warn;
    std::lock_guard<std::mutex> lk(s->mu);
    ++s->st.graph_captures;
  } else {
    e.exec = nullptr;
    e.broken = rc == SD_EUNSUPPORTED;  // a path with host decisions: never captured
  }
  if (g) (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  sd::clear_error();
  (void)nc;
  return SD_OK;
}

// One search of the pieces' columns, concatenated in order, against Y on the
// slot's context; each piece's results go to its own request.  A query's
// result does not depend on the other columns of the search, so packing
// changes no answer.  Items of at most kGraphCols columns replay a captured
// graph when they can.
int run_segs(sd_server* s, int slot, sd_context* ctx, sd_matrix* X, const std::vector<Seg>& segs) {
  const Request* r0 = segs[0].r;
  int64_t nc = 0, nnz = 0;
  for (const Seg& g : segs) {
    nc += g.c1 - g.c0;
    nnz += g.r->colptr[g.c1] - g.r->colptr[g.c0];
  }
  if (graph_eligible(s, r0->knn, nc)) {
    bool done = false;
    const int grc = guarded_call([&] { return run_graph_item(s, slot, segs, &done); });
    if (grc == SD_OK && done) return SD_OK;
    sd::clear_error();
  }
  std::vector<int64_t> p((size_t)nc + 1);
  std::vector<int32_t> ri;
  std::vector<double> vx;
  const int32_t* rows = nullptr;
  const double* vals = nullptr;
  if (segs.size() == 1) {  // one request's columns: its arrays are contiguous
    const int64_t base = r0->colptr[segs[0].c0];
    for (int64_t j = 0; j <= nc; ++j) p[j] = r0->colptr[segs[0].c0 + j] - base;
    rows = r0->rowind + base;
    vals = r0->values + base;
  } else {
    ri.resize((size_t)nnz);
    vx.resize((size_t)nnz);
    int64_t j = 0, e = 0;
    for (const Seg& g : segs) {
      const int64_t b0 = g.r->colptr[g.c0], b1 = g.r->colptr[g.c1];
      std::copy(g.r->rowind + b0, g.r->rowind + b1, ri.begin() + e);
      std::copy(g.r->values + b0, g.r->values + b1, vx.begin() + e);
      for (int64_t c = g.c0; c < g.c1; ++c) p[(size_t)j++] = e + g.r->colptr[c] - b0;
      e += b1 - b0;
    }
    p[(size_t)nc] = e;
    rows = ri.data();
    vals = vx.data();
  }
  SD_TRY(sd::matrix_load(ctx, X, r0->nrow, nc, p.data(), rows, vals));
  const int32_t k = r0->knn.k;
  std::vector<int32_t> idx((size_t)nc * k);
  std::vector<double> dist((size_t)nc * k);
  int32_t keff = 0, warn = 0;
  sd_knn_params prm = r0->knn;
  prm.include_self = 0;
  SD_TRY(sd_knn(ctx, X, s->Y, &prm, idx.data(), dist.data(), &keff, &warn));
  scatter_results(s, segs, idx.data(), dist.data(), nc, keff, warn);
  return SD_OK;
}

// Shared-neighbour graph of the query columns: two queries are joined when
// their neighbour sets in Y intersect, weighted by the Jaccard index of the
// sets and kept when the weight exceeds prune (the "shared" edge rule of
// R/sparseSNN.R:206-218 with neighbour ids in Y's column space).
int run_snn(sd_server* s, sd_context* ctx, Request* r) {
  const int64_t n = r->ncol;
  const int32_t k = r->k_eff;
  if (n >= ((int64_t)1 << 31) || s->Y->ncol >= ((int64_t)1 << 31) || k < 1)
    return sd::set_error(SD_EINVAL, "SNN needs fewer than 2^31 columns and k >= 1");
  sd::DevBuf& di = ctx->snn_scratch[0];
  SD_TRY(di.ensure(sizeof(int32_t) * (size_t)n * (size_t)k));
  SD_HIP(hipMemcpyAsync(di.ptr, r->idx.data(), sizeof(int32_t) * (size_t)n * (size_t)k,
                        hipMemcpyHostToDevice, ctx->stream));
  auto* g = new sd_graph();
  g->device = ctx->device;
  const int rc = sd::snn_build(ctx, di.as<int32_t>(), (int)n, k, 0, r->prune, 0, 1, g, nullptr,
                               (int)s->Y->ncol);
  if (rc != SD_OK) {
    delete g;
    return rc;
  }
  r->graph = g;
  return SD_OK;
}

// The compatible admitted request, other than those in segs, with the fewest
// unscheduled columns (ties: earliest).
Request* pick_compatible_locked(sd_server* s, const std::vector<Seg>& segs) {
  Request* best = nullptr;
  for (Request* r : s->active) {
    if (r->next_col >= r->ncol || r->status != SD_OK) continue;
    if (!same_search(r->knn, segs[0].r->knn)) continue;
    bool taken = false;
    for (const Seg& g : segs) taken = taken || g.r == r;
    if (taken) continue;
    const int64_t left = r->ncol - r->next_col;
    if (!best || left < best->ncol - best->next_col ||
        (left == best->ncol - best->next_col && r->ticket < best->ticket))
      best = r;
  }
  return best;
}

void worker(sd_server* s, int slot) {
  sd_context* ctx = s->ctx[slot];
  sd_matrix* X = s->qmat[slot];
  (void)hipSetDevice(s->device);
  constexpr int64_t kTile = 256;  // query columns of one slab tile
  for (;;) {
    std::vector<Seg> segs;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || (!s->paused && pick_locked(s) != nullptr); });
      if (s->stop) return;
      Request* r = pick_locked(s);
      const int64_t c0 = r->next_col, c1 = std::min(r->ncol, c0 + s->chunk);
      r->next_col = c1;
      ++r->inflight;
      segs.push_back(Seg{r, c0, c1});
      // Continuous batching: a slab computes whole tiles of 256 query
      // columns, so columns of other waiting requests with the same search
      // fill the last tile at no extra slab work.
      int64_t total = c1 - c0;
      const int64_t target = std::min(s->chunk, (total + kTile - 1) / kTile * kTile);
      while (total < target) {
        Request* o = pick_compatible_locked(s, segs);
        if (!o) break;
        const int64_t take = std::min(o->ncol - o->next_col, target - total);
        segs.push_back(Seg{o, o->next_col, o->next_col + take});
        o->next_col += take;
        ++o->inflight;
        total += take;
      }
      if (segs.size() > 1) ++s->st.packed_items;
    }
    // a failed packed search is run again piece by piece, so that an error
    // (say, non-finite input of one request) is charged to its own request
    std::vector<int> rcs(segs.size(), SD_OK);
    std::vector<std::string> msgs(segs.size());
    int rc = guarded_call([&] { return run_segs(s, slot, ctx, X, segs); });
    if (rc != SD_OK && rc != SD_ECANCELLED && segs.size() > 1) {
      for (size_t i = 0; i < segs.size(); ++i) {
        rcs[i] = guarded_call([&] { return run_segs(s, slot, ctx, X, {segs[i]}); });
        if (rcs[i] != SD_OK) msgs[i] = sd_last_error();
      }
    } else if (rc != SD_OK) {
      for (size_t i = 0; i < segs.size(); ++i) {
        rcs[i] = rc;
        msgs[i] = sd_last_error();
      }
    }
    for (size_t i = 0; i < segs.size(); ++i) {
      Request* r = segs[i].r;
      bool complete = false, need_snn = false;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        --r->inflight;
        if (rcs[i] == SD_OK) {
          r->done_cols += segs[i].c1 - segs[i].c0;
        } else if (r->status == SD_OK) {
          r->status = rcs[i];
          r->error = msgs[i];
          r->next_col = r->ncol;  // schedule nothing more of it
        }
        // the item that returns last, with nothing left to schedule, completes it
        complete = r->inflight == 0 && r->next_col == r->ncol;
        need_snn = complete && r->status == SD_OK && r->snn;
        if (complete && !need_snn) finish_locked(s, r);
      }
      if (need_snn) {
        const int src = guarded_call([&] { return run_snn(s, ctx, r); });
        std::lock_guard<std::mutex> lk(s->mu);
        if (src != SD_OK) {
          r->status = src;
          r->error = sd_last_error();
        }
        finish_locked(s, r);
      }
    }
  }
}

Request* find_locked(sd_server* s, int64_t ticket) {
  auto it = s->all.find(ticket);
  return it == s->all.end() ? nullptr : it->second.get();
}

}  // namespace

extern "C" {

int sd_server_create(sd_context* ctx, const sd_matrix* Y, const sd_server_params* params,
                     sd_server** out) {
  return guarded_call([&]() -> int {
    if (!ctx || !Y || !out) return sd::set_error(SD_EINVAL, "NULL argument");
    sd_server_params p{};
    if (params) p = *params;
    if (p.slots < 0 || p.slots > 8) return sd::set_error(SD_EINVAL, "slots must be in 0..8");
    if (p.chunk_cols < 0) return sd::set_error(SD_EINVAL, "chunk_cols must be >= 0");
    if (p.budget_bytes < 0 || p.max_pending < 0)
      return sd::set_error(SD_EINVAL, "budget_bytes and max_pending must be >= 0");
    if (Y->ncol < 1) return sd::set_error(SD_EINVAL, "'Y' must have at least one column.");
    auto s = std::make_unique<sd_server>();
    s->device = ctx->device;
    s->Y = Y;
    s->slots = p.slots ? p.slots : 4;
    s->chunk = p.chunk_cols ? p.chunk_cols : 4096;
    s->budget = p.budget_bytes ? p.budget_bytes : ((int64_t)16 << 30);
    s->max_pending = p.max_pending;
    s->paused = p.paused != 0;
    auto undo = [&] {
      for (sd_matrix* m : s->qmat) sd_matrix_free(m);
      for (sd_context* d : s->ctx) sd_context_destroy(d);
      for (GraphSlot& gs : s->gslot) {
        if (gs.X) sd_matrix_free(gs.X);
        if (gs.ctx) sd_context_destroy(gs.ctx);
      }
    };
    for (int i = 0; i < s->slots; ++i) {
      sd_context* c = nullptr;
      int rc = sd_context_create(s->device, &c);
      if (rc == SD_OK) {
        s->ctx.push_back(c);
        rc = sd::context_use_pool(c);
      }
      if (rc != SD_OK) {
        undo();
        return rc;
      }
      auto* m = new sd_matrix();
      m->ctx = c;
      m->device = c->device;
      sd::matrix_use_pool(m, c->stream);
      s->qmat.push_back(m);
    }
    if (p.graphs) {
      // each slot's graph items run on a context of their own, with the
      // fixed threshold-sample rule (a tuned choice could change buffer sizes
      // between the ordinary run and the capture)
      s->gslot.resize((size_t)s->slots);
      for (GraphSlot& gs : s->gslot) {
        int rc = sd_context_create(s->device, &gs.ctx);
        if (rc == SD_OK) rc = sd::context_use_pool(gs.ctx);
        if (rc != SD_OK) {
          undo();
          return rc;
        }
        gs.ctx->sample_tuning = 0;
        gs.X = new sd_matrix();
        gs.X->ctx = gs.ctx;
        gs.X->device = gs.ctx->device;
        sd::matrix_use_pool(gs.X, gs.ctx->stream);
      }
    }
    for (int i = 0; i < s->slots; ++i) s->workers.emplace_back(worker, s.get(), i);
    *out = s.release();
    return SD_OK;
  });
}

int sd_server_submit(sd_server* s, const sd_request* req, int64_t* ticket) {
  return guarded_call([&]() -> int {
    if (!s || !req || !ticket) return sd::set_error(SD_EINVAL, "NULL argument");
    if (req->ncol < 1) return sd::set_error(SD_EINVAL, "'X' must have at least one column.");
    if (req->nrow != s->Y->nrow)
      return sd::set_error(SD_EINVAL, "'X' and 'Y' must have the same number of rows.");
    if (!req->colptr || (req->colptr[req->ncol] > 0 && (!req->rowind || !req->values)))
      return sd::set_error(SD_EINVAL, "NULL argument");
    if (req->colptr[0] != 0) return sd::set_error(SD_EINVAL, "colptr[0] must be 0");
    if (req->knn.k < 1)
      return sd::set_error(SD_EINVAL,
                           "'k' must be a positive integer no larger than .Machine$integer.max.");
    if (req->snn && !(req->prune >= 0.0 && req->prune <= 1.0))
      return sd::set_error(SD_EINVAL, "'prune' must be a finite number between 0 and 1.");
    {
      // refuse before copying anything when the queue is full
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->max_pending > 0 && (int64_t)s->pending.size() >= s->max_pending) {
        ++s->st.refused;
        return sd::set_error(SD_EBUDGET, "server queue full (%lld requests waiting)",
                             (long long)s->pending.size());
      }
    }
    auto r = std::make_unique<Request>();
    r->nrow = req->nrow;
    r->ncol = req->ncol;
    r->nnz = req->colptr[req->ncol];
    if (req->borrow) {
      r->colptr = req->colptr;
      r->rowind = req->rowind;
      r->values = req->values;
    } else {
      r->own_p.assign(req->colptr, req->colptr + req->ncol + 1);
      if (r->nnz > 0) {
        r->own_i.assign(req->rowind, req->rowind + r->nnz);
        r->own_x.assign(req->values, req->values + r->nnz);
      }
      r->colptr = r->own_p.data();
      r->rowind = r->own_i.data();
      r->values = r->own_x.data();
    }
    r->knn = req->knn;
    r->snn = req->snn;
    r->prune = req->prune;
    r->reserve = reserve_estimate(s, *r);
    if (r->reserve > s->budget)
      return sd::set_error(SD_EBUDGET,
                           "request needs an estimated %lld bytes of device memory, more than "
                           "the server budget of %lld",
                           (long long)r->reserve, (long long)s->budget);
    r->idx.assign((size_t)r->ncol * (size_t)r->knn.k, sd::kNaInt);
    r->dist.assign((size_t)r->ncol * (size_t)r->knn.k, 0.0);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->max_pending > 0 && (int64_t)s->pending.size() >= s->max_pending) {
      ++s->st.refused;
      return sd::set_error(SD_EBUDGET, "server queue full (%lld requests waiting)",
                           (long long)s->pending.size());
    }
    r->ticket = s->next_ticket++;
    r->t_submit = Clock::now();
    *ticket = r->ticket;
    s->pending.push_back(r.get());
    s->all.emplace(r->ticket, std::move(r));
    ++s->st.submitted;
    admit_locked(s);
    s->cv_work.notify_all();
    return SD_OK;
  });
}

int sd_server_wait(sd_server* s, int64_t ticket, double timeout_s, sd_request_info* info) {
  return guarded_call([&]() -> int {
    if (!s || !info) return sd::set_error(SD_EINVAL, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    Request* r = find_locked(s, ticket);
    if (!r) return sd::set_error(SD_EINVAL, "unknown ticket %lld", (long long)ticket);
    auto ready = [&] { return r->finished; };
    // sd_server_destroy finishes every request and then waits until no
    // caller is left in here before it frees the server
    ++s->waiters;
    struct Leave {
      sd_server* s;
      ~Leave() {
        if (--s->waiters == 0) s->cv_gone.notify_all();
      }
    } leave{s};
    if (timeout_s < 0)
      s->cv_done.wait(lk, ready);
    else
      s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
    *info = sd_request_info{};
    info->done = r->finished ? 1 : 0;
    info->status = r->status;
    info->k_eff = r->k_eff;
    info->warn = r->warn;
    info->finish_order = r->finish_order;
    if (r->admitted) info->queued_ms = ms_between(r->t_submit, r->t_admit);
    if (r->finished) {
      info->run_ms = ms_between(r->t_admit, r->t_done);
      info->total_ms = ms_between(r->t_submit, r->t_done);
    }
    return SD_OK;
  });
}

const char* sd_server_error(sd_server* s, int64_t ticket) {
  if (!s) return "";
  std::lock_guard<std::mutex> lk(s->mu);
  Request* r = find_locked(s, ticket);
  s->last_msg = r ? r->error : std::string();
  return s->last_msg.c_str();
}

int sd_server_result(sd_server* s, int64_t ticket, int32_t* idx, double* dist, sd_graph** snn) {
  return guarded_call([&]() -> int {
    if (!s || !idx || !dist) return sd::set_error(SD_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    Request* r = find_locked(s, ticket);
    if (!r) return sd::set_error(SD_EINVAL, "unknown ticket %lld", (long long)ticket);
    if (!r->finished) return sd::set_error(SD_EINVAL, "request %lld has not finished",
                                           (long long)ticket);
    if (r->status != SD_OK) return sd::set_error(r->status, "%s", r->error.c_str());
    const size_t m = (size_t)r->ncol * (size_t)r->k_eff;
    std::copy(r->idx.begin(), r->idx.begin() + m, idx);
    std::copy(r->dist.begin(), r->dist.begin() + m, dist);
    if (snn) {
      *snn = r->graph;
      r->graph = nullptr;
    }
    return SD_OK;
  });
}

int sd_server_release(sd_server* s, int64_t ticket) {
  return guarded_call([&]() -> int {
    if (!s) return sd::set_error(SD_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    Request* r = find_locked(s, ticket);
    if (!r) return sd::set_error(SD_EINVAL, "unknown ticket %lld", (long long)ticket);
    if (!r->finished) return sd::set_error(SD_EINVAL, "request %lld has not finished",
                                           (long long)ticket);
    if (r->graph) sd_graph_free(r->graph);
    s->all.erase(ticket);
    return SD_OK;
  });
}

int sd_server_stats_get(sd_server* s, sd_server_stats* out) {
  if (!s || !out) return sd::set_error(SD_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->pending = (int64_t)s->pending.size();
  out->active = (int64_t)s->active.size();
  out->reserved_bytes = s->reserved + s->reserved_lane;
  out->waiting_callers = s->waiters;
  return SD_OK;
}

int sd_server_resume(sd_server* s) {
  if (!s) return sd::set_error(SD_EINVAL, "NULL argument");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->paused = false;
  }
  s->cv_work.notify_all();
  return SD_OK;
}

int sd_server_destroy(sd_server* s) {
  if (!s) return SD_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  for (sd_context* c : s->ctx) sd_cancel(c);
  for (GraphSlot& gs : s->gslot) sd_cancel(gs.ctx);
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  {
    std::unique_lock<std::mutex> lk(s->mu);
    for (auto& kv : s->all) {
      Request* r = kv.second.get();
      if (!r->finished) {
        r->finished = true;
        r->finish_order = ++s->finished_count;
        r->t_done = Clock::now();
        if (r->status == SD_OK) {
          r->status = SD_ECANCELLED;
          r->error = "Computation interrupted by the user.";
        }
      }
    }
    s->cv_done.notify_all();
    // callers blocked in sd_server_wait read their request once more
    s->cv_gone.wait(lk, [&] { return s->waiters == 0; });
    for (auto& kv : s->all)
      if (kv.second->graph) sd_graph_free(kv.second->graph);
  }
  for (sd_matrix* m : s->qmat) sd_matrix_free(m);
  for (sd_context* c : s->ctx) sd_context_destroy(c);
  for (GraphSlot& gs : s->gslot) {
    for (auto& kv : gs.graphs)
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    sd_matrix_free(gs.X);
    sd_context_destroy(gs.ctx);
  }
  delete s;
  return SD_OK;
}

}  // extern "C"
