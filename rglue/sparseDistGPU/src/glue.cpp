// glue.cpp -- Rcpp bindings of libsd_hip.so's C ABI (include/sd.h).
//
// Only plain pointers cross into the library.  dgCMatrix slots are passed as
// they are (colptr widened to int64); results are written straight into
// IntegerMatrix / NumericMatrix storage, whose NA_integer_ / NA_real_
// padding the library already produces.  Long calls run on a worker thread
// while the R thread polls for interrupts and forwards them with sd_cancel.
#include <Rcpp.h>

#include <atomic>
#include <memory>
#include <chrono>
#include <thread>
#include <vector>

#include "sd.h"

namespace {

sd_context* ctx() {
  static sd_context* c = nullptr;
  if (!c && sd_context_create(0, &c) != SD_OK) Rcpp::stop(sd_last_error());
  return c;
}

void check(int rc, const char* msg = nullptr) {
  if (rc != SD_OK) Rcpp::stop(msg ? msg : sd_last_error());
}

// A device copy of a dgCMatrix, freed with the object.
struct DevMat {
  sd_matrix* m = nullptr;
  explicit DevMat(Rcpp::S4 M) {
    Rcpp::IntegerVector dim = M.slot("Dim"), p = M.slot("p"), i = M.slot("i");
    Rcpp::NumericVector x = M.slot("x");
    std::vector<int64_t> p64(p.begin(), p.end());
    check(sd_matrix_upload_csc(ctx(), dim[0], dim[1], p64.data(), i.begin(), x.begin(), &m));
  }
  ~DevMat() {
    if (m) sd_matrix_free(m);
  }
};

void check_interrupt(void*) { R_CheckUserInterrupt(); }

// Runs f() on a worker thread; an R interrupt cancels it.
template <class F>
int interruptible(F f) {
  std::atomic<bool> done{false};
  int rc = SD_OK;
  std::string err;
  std::thread t([&] {
    rc = f();
    if (rc != SD_OK) err = sd_last_error();  // the message is per thread
    done = true;
  });
  while (!done) {
    if (!R_ToplevelExec(check_interrupt, nullptr)) sd_cancel(ctx());
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
  }
  t.join();
  if (rc == SD_ECANCELLED) Rcpp::stop("Computation interrupted by the user.");
  if (rc != SD_OK) Rcpp::stop(err);
  return rc;
}

sd_knn_params params(int k, int method, bool dist, bool include_self, bool drop_degenerate,
                     int block_size) {
  sd_knn_params p{};
  p.k = k;
  p.method = method;
  p.dist = dist;
  p.include_self = include_self;
  p.drop_degenerate = drop_degenerate;
  p.precision = SD_PREC_AUTO;
  p.block_size = block_size;
  return p;
}

}  // namespace

// [[Rcpp::export(.knn_gpu)]]
Rcpp::List knn_gpu(Rcpp::S4 X, Rcpp::Nullable<Rcpp::S4> Y, int k, int method, bool dist,
                   bool include_self, bool drop_degenerate, int block_size) {
  DevMat dx(X);
  std::unique_ptr<DevMat> dy;
  if (Y.isNotNull()) dy.reset(new DevMat(Rcpp::S4(Y)));
  const sd_knn_params p = params(k, method, dist, include_self, drop_degenerate, block_size);
  const int nq = Rcpp::IntegerVector(X.slot("Dim"))[1];
  Rcpp::IntegerVector idx((R_xlen_t)nq * k);
  Rcpp::NumericVector dst((R_xlen_t)nq * k);
  int32_t keff = 0, warn = 0;
  interruptible([&] {
    return sd_knn(ctx(), dx.m, dy ? dy->m : nullptr, &p, idx.begin(), dst.begin(), &keff, &warn);
  });
  Rcpp::IntegerMatrix I(nq, keff);
  Rcpp::NumericMatrix D(nq, keff);
  std::copy(idx.begin(), idx.begin() + (R_xlen_t)nq * keff, I.begin());
  std::copy(dst.begin(), dst.begin() + (R_xlen_t)nq * keff, D.begin());
  return Rcpp::List::create(Rcpp::Named("idx") = I, Rcpp::Named("dist") = D,
                            Rcpp::Named("k") = keff, Rcpp::Named("warn") = warn);
}

// [[Rcpp::export(.knn_cross_gpu)]]
Rcpp::List knn_cross_gpu(Rcpp::S4 X, Rcpp::S4 Y, int k, int method, bool dist,
                         bool drop_degenerate, int block_size) {
  DevMat dx(X), dy(Y);
  const sd_knn_params p = params(k, method, dist, false, drop_degenerate, block_size);
  const int nq = Rcpp::IntegerVector(X.slot("Dim"))[1];
  const int nr = Rcpp::IntegerVector(Y.slot("Dim"))[1];
  Rcpp::IntegerVector fi((R_xlen_t)nq * k), ri((R_xlen_t)nr * k);
  Rcpp::NumericVector fd((R_xlen_t)nq * k), rd((R_xlen_t)nr * k);
  int32_t kf = 0, kr = 0, warn = 0;
  interruptible([&] {
    return sd_knn_cross(ctx(), dx.m, dy.m, &p, fi.begin(), fd.begin(), &kf, ri.begin(),
                        rd.begin(), &kr, &warn);
  });
  auto pack = [](int n, int ke, Rcpp::IntegerVector& i, Rcpp::NumericVector& d) {
    Rcpp::IntegerMatrix I(n, ke);
    Rcpp::NumericMatrix D(n, ke);
    std::copy(i.begin(), i.begin() + (R_xlen_t)n * ke, I.begin());
    std::copy(d.begin(), d.begin() + (R_xlen_t)n * ke, D.begin());
    return Rcpp::List::create(Rcpp::Named("idx") = I, Rcpp::Named("dist") = D);
  };
  return Rcpp::List::create(Rcpp::Named("forward") = pack(nq, kf, fi, fd),
                            Rcpp::Named("reverse") = pack(nr, kr, ri, rd),
                            Rcpp::Named("k_forward") = kf, Rcpp::Named("k_reverse") = kr,
                            Rcpp::Named("warn") = warn);
}

namespace {

Rcpp::S4 graph_to_dgc(sd_graph* g) {
  int64_t n = 0, nnz = 0;
  sd_graph_info(g, &n, &nnz);
  std::vector<int64_t> p((size_t)n + 1);
  Rcpp::IntegerVector i((R_xlen_t)nnz);
  Rcpp::NumericVector x((R_xlen_t)nnz);
  const int rc = sd_graph_copy(g, p.data(), i.begin(), x.begin());
  sd_graph_free(g);
  check(rc);
  Rcpp::S4 out("dgCMatrix");
  out.slot("p") = Rcpp::IntegerVector(p.begin(), p.end());
  out.slot("i") = i;
  out.slot("x") = x;
  out.slot("Dim") = Rcpp::IntegerVector::create((int)n, (int)n);
  return out;
}

}  // namespace

// [[Rcpp::export(.snn_gpu)]]
Rcpp::S4 snn_gpu(Rcpp::IntegerMatrix idx, bool include_self, double prune, bool symmetrise,
                 int edges) {
  sd_graph* g = nullptr;
  check(sd_snn(ctx(), idx.begin(), idx.nrow(), idx.ncol(), include_self, prune, symmetrise,
               edges, 0, &g));
  return graph_to_dgc(g);
}

// ---- serving -----------------------------------------------------------------
namespace {

struct Server {
  sd_server* s = nullptr;
  DevMat* Y = nullptr;
  ~Server() {
    if (s) sd_server_destroy(s);
    delete Y;
  }
};

void server_finalizer(SEXP p) {
  delete static_cast<Server*>(R_ExternalPtrAddr(p));
  R_ClearExternalPtr(p);
}

Server* server_of(SEXP p) {
  auto* s = static_cast<Server*>(R_ExternalPtrAddr(p));
  if (!s) Rcpp::stop("the server has been closed");
  return s;
}

}  // namespace

// [[Rcpp::export(.server_create)]]
SEXP server_create(Rcpp::S4 Y, int slots, int chunk_cols, double budget_bytes,
                   double max_pending) {
  auto* s = new Server();
  s->Y = new DevMat(Y);
  sd_server_params p{slots, chunk_cols, (int64_t)budget_bytes, (int64_t)max_pending};
  const int rc = sd_server_create(ctx(), s->Y->m, &p, &s->s);
  if (rc != SD_OK) {
    delete s;
    check(rc);
  }
  SEXP ptr = PROTECT(R_MakeExternalPtr(s, R_NilValue, R_NilValue));
  R_RegisterCFinalizerEx(ptr, server_finalizer, TRUE);
  UNPROTECT(1);
  return ptr;
}

// [[Rcpp::export(.server_submit)]]
double server_submit(SEXP srv, Rcpp::S4 X, int k, int method, bool dist, bool drop_degenerate,
                     bool snn, double prune) {
  Server* s = server_of(srv);
  Rcpp::IntegerVector dim = X.slot("Dim"), p = X.slot("p"), i = X.slot("i");
  Rcpp::NumericVector x = X.slot("x");
  std::vector<int64_t> p64(p.begin(), p.end());
  sd_request r{};
  r.nrow = dim[0];
  r.ncol = dim[1];
  r.colptr = p64.data();
  r.rowind = i.begin();
  r.values = x.begin();
  r.knn = params(k, method, dist, false, drop_degenerate, 256);
  r.snn = snn;
  r.borrow = 0;  // the server copies X: R may move or free it afterwards
  r.prune = prune;
  int64_t ticket = 0;
  check(sd_server_submit(s->s, &r, &ticket));
  return (double)ticket;
}

// [[Rcpp::export(.server_result)]]
Rcpp::List server_result(SEXP srv, double ticket, int ncol) {
  Server* s = server_of(srv);
  sd_request_info info{};
  interruptible([&] { return sd_server_wait(s->s, (int64_t)ticket, -1.0, &info); });
  if (info.status != SD_OK) {
    const std::string msg = sd_server_error(s->s, (int64_t)ticket);
    sd_server_release(s->s, (int64_t)ticket);
    Rcpp::stop(msg);
  }
  Rcpp::IntegerMatrix I(ncol, info.k_eff);
  Rcpp::NumericMatrix D(ncol, info.k_eff);
  sd_graph* g = nullptr;
  check(sd_server_result(s->s, (int64_t)ticket, I.begin(), D.begin(), &g));
  sd_server_release(s->s, (int64_t)ticket);
  Rcpp::List out = Rcpp::List::create(
      Rcpp::Named("idx") = I, Rcpp::Named("dist") = D, Rcpp::Named("warn") = info.warn,
      Rcpp::Named("queued_ms") = info.queued_ms, Rcpp::Named("total_ms") = info.total_ms);
  if (g) out["snn"] = graph_to_dgc(g);
  return out;
}

// [[Rcpp::export(.server_close)]]
void server_close(SEXP srv) {
  delete static_cast<Server*>(R_ExternalPtrAddr(srv));
  R_ClearExternalPtr(srv);
}

// [[Rcpp::export(.autotune)]]
Rcpp::List autotune(bool measure) {
  sd_tune_info t{};
  check(sd_autotune(ctx(), measure, &t));
  return Rcpp::List::create(
      Rcpp::Named("bf16_ops") = t.bf16_ops, Rcpp::Named("fp4_ops") = t.fp4_ops,
      Rcpp::Named("merge_steps") = t.merge_steps, Rcpp::Named("bf16_fixed_ms") = t.bf16_fixed_ms,
      Rcpp::Named("fp4_fixed_ms") = t.fp4_fixed_ms,
      Rcpp::Named("exact_fixed_ms") = t.exact_fixed_ms,
      Rcpp::Named("step_latency_us") = t.step_latency_us,
      Rcpp::Named("calibrated") = (bool)t.calibrated);
}
