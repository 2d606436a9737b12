// sd_tune.cpp -- the path chooser's cost model and its on-device calibration.
//
// A search of nq query columns against nref reference columns with r rows
// costs, on the densified MFMA path, a fixed overhead plus whole query tiles
// of 256 columns times 2 r nref dense-equivalent operations at the filtered
// pass's rate; on the FP64 merge walk it costs a smaller overhead plus
// nq nref (nnz_q + nnz_ref) merge steps at the walk's throughput, but no less
// than one pair's nnz_q + nnz_ref sequential steps (a lone query's walks run
// side by side, each at the latency of a dependent load per step).  The
// densified path of the re-ranked methods adds that floor for its FP64
// re-rank: one candidate's column per lane (cosine walks the reference
// column against a bitmap of the query; euclidean and pearson walk both).  A
// self-search large enough for the filtered pass scores each pair once, half
// the slab.  Both paths are exact, so the choice
// changes only the time.  sd_autotune replaces the default rates by ones
// measured on synthetic inputs of known size.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "sd_internal.h"

namespace sd {

// ---- threshold-sample tuner ------------------------------------------------
// The filtered search's sample (every stride-th reference tile at rank Lp)
// costs the pre-pass 1/stride of the pairs and leaves about stride x Lp
// candidates per query for the filtered pass's epilogue and the selection.
// Which balance is fastest depends on the row length (a long K amortises the
// epilogue), the density of the layout's values, the list length and the
// symmetry of the pass, so each class of searches times its candidates: the
// first searches of a class large enough to time (at least 2^28 scored
// pairs) run the candidates not yet timed, one each (a search that had to
// allocate device memory is not counted); later ones run the fastest -- the
// fixed rule's choice unless another is at least 3% faster -- and keep
// refining its time.  Every candidate returns the same
// results, so exploring changes only the time.  The entries are process-wide
// (all contexts share what they measured) and listed by sd_tune_samples.
SampleChoice sample_candidate(int c, int L) {
  SampleChoice s;
  switch (c) {
    case 0: s = {32, L}; break;
    case 1: s = {64, fused_list_len((L + 1) / 2)}; break;
    case 2: s = {128, fused_list_len((L + 3) / 4)}; break;
    default: s = {64, fused_list_len(std::max((L + 3) / 4, 8))}; break;
  }
  return s;
}

bool SampleKey::operator<(const SampleKey& o) const {
  const int a[6] = {method, kind, density, row_bytes, L, tri};
  const int b[6] = {o.method, o.kind, o.density, o.row_bytes, o.L, o.tri};
  return std::lexicographical_compare(a, a + 6, b, b + 6);
}

namespace {
struct SampleEntry {
  double sec_per_pair[kSampleCandidates] = {0, 0, 0, 0};
  int64_t runs[kSampleCandidates] = {0, 0, 0, 0};
  int last_L = 0;     // list length of the last timed search (for the listing)
  int last_rule = -1; // the fixed rule's candidate and the one run, last search
  int last_run = -1;
};
std::mutex g_sample_mu;
std::map<SampleKey, SampleEntry>& sample_table() {
  static std::map<SampleKey, SampleEntry> t;
  return t;
}
}  // namespace

int sample_plan(const SampleKey& key, const bool allowed[kSampleCandidates], int def, bool tuning,
                bool timeable, bool* record) {
  *record = false;
  if (!tuning) return def;
  std::lock_guard<std::mutex> lk(g_sample_mu);
  auto it = sample_table().find(key);
  const SampleEntry* e = it == sample_table().end() ? nullptr : &it->second;
  if (timeable) {
    // untimed candidates first, the default among them first
    if (allowed[def] && (!e || e->runs[def] == 0)) {
      *record = true;
      return def;
    }
    for (int c = 0; c < kSampleCandidates; ++c)
      if (allowed[c] && (!e || e->runs[c] == 0)) {
        *record = true;
        return c;
      }
  }
  if (!e) return def;
  int best = -1;
  for (int c = 0; c < kSampleCandidates; ++c)
    if (allowed[c] && e->runs[c] > 0 && (best < 0 || e->sec_per_pair[c] < e->sec_per_pair[best]))
      best = c;
  if (best < 0) return def;
  // the rule's choice stays unless another candidate is clearly faster
  // (timing noise between single searches is a few percent)
  if (best != def && allowed[def] && e->runs[def] > 0 &&
      e->sec_per_pair[best] > 0.97 * e->sec_per_pair[def])
    best = def;
  *record = timeable;
  return best;
}

void sample_note(const SampleKey& key, int rule, int ran) {
  std::lock_guard<std::mutex> lk(g_sample_mu);
  auto it = sample_table().find(key);
  if (it == sample_table().end()) return;
  it->second.last_rule = rule;
  it->second.last_run = ran;
}

void sample_record(const SampleKey& key, int cand, int L, double sec_per_pair) {
  if (cand < 0 || cand >= kSampleCandidates || !(sec_per_pair > 0)) return;
  std::lock_guard<std::mutex> lk(g_sample_mu);
  SampleEntry& e = sample_table()[key];
  e.last_L = L;
  // the first time stands; later ones move it a quarter of the way
  e.sec_per_pair[cand] = e.runs[cand] == 0 ? sec_per_pair
                                           : 0.75 * e.sec_per_pair[cand] + 0.25 * sec_per_pair;
  ++e.runs[cand];
}

bool tune_prefer_exact(const sd_context* ctx, int kind, int64_t nq, int64_t nref, int64_t kpad,
                       double nnz_q, double nnz_ref, bool symmetric) {
  const TuneRates& t = ctx->tune;
  const bool bf16 = kind == kDenseBf16Unit || kind == kDenseBf16Split || kind == kDenseBf16Raw;
  const double rate = kind == kDenseF4Pattern ? t.fp4_ops : bf16 ? t.bf16_ops : t.fp32_ops;
  const double fixed = kind == kDenseF4Pattern ? t.fp4_fixed_s : bf16 ? t.bf16_fixed_s : t.fp32_fixed_s;
  // the split layout runs three products per row
  const double kwork = (double)kpad * (kind == kDenseBf16Split ? 3.0 : 1.0);
  const double qcols = (double)((nq + 255) / 256 * 256);
  const bool pattern = kind == kDenseF4Pattern || kind == kDenseF32Pattern;
  const bool unit = kind == kDenseBf16Unit || kind == kDenseF32Unit;
  const double rerank = pattern ? 0.0 : (unit ? nnz_ref : nnz_q + nnz_ref) * t.step_latency_s;
  // a self-search's filtered pass scores each unordered pair once
  const double slab = qcols * (double)nref * 2.0 * kwork / rate * (symmetric ? 0.5 : 1.0);
  const double fused = fixed + slab + rerank;
  const double exact =
      t.exact_fixed_s + std::max((double)nq * (double)nref * (nnz_q + nnz_ref) / t.merge_steps,
                                 (nnz_q + nnz_ref) * t.step_latency_s);
  return exact < fused;
}

namespace {

// CSC with `per` distinct random rows per column (partial Fisher-Yates over a
// running permutation, then sorted) and values in [0.5, 1.5).
struct SynthCsc {
  std::vector<int64_t> p;
  std::vector<int32_t> i;
  std::vector<double> x;
};

SynthCsc synth_csc(int64_t nrow, int64_t ncol, int per, uint64_t seed) {
  SynthCsc m;
  m.p.resize((size_t)ncol + 1);
  m.i.resize((size_t)ncol * per);
  m.x.resize((size_t)ncol * per);
  std::vector<int32_t> perm((size_t)nrow);
  for (int64_t r = 0; r < nrow; ++r) perm[(size_t)r] = (int32_t)r;
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
  auto next = [&]() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  };
  for (int64_t j = 0; j < ncol; ++j) {
    m.p[(size_t)j] = j * per;
    for (int t = 0; t < per; ++t) {
      const int64_t u = t + (int64_t)(next() % (uint64_t)(nrow - t));
      std::swap(perm[(size_t)t], perm[(size_t)u]);
    }
    int32_t* rows = m.i.data() + j * per;
    std::copy(perm.begin(), perm.begin() + per, rows);
    std::sort(rows, rows + per);
    for (int t = 0; t < per; ++t)
      m.x[(size_t)(j * per + t)] = 0.5 + (double)(next() >> 11) * (1.0 / 9007199254740992.0);
  }
  m.p[(size_t)ncol] = ncol * per;
  return m;
}

int upload(sd_context* ctx, int64_t nrow, const SynthCsc& m, sd_matrix** out) {
  return sd_matrix_upload_csc(ctx, nrow, (int64_t)m.p.size() - 1, m.p.data(), m.i.data(),
                              m.x.data(), out);
}

// Wall time of one sd_knn call (synchronous: results are copied to the host).
int time_knn(sd_context* ctx, sd_matrix* Q, sd_matrix* Y, int method, int prec, double* sec) {
  sd_knn_params p{};
  p.k = 10;
  p.method = method;
  p.dist = 1;
  p.drop_degenerate = 1;
  p.precision = prec;
  p.block_size = 4096;
  std::vector<int32_t> idx((size_t)Q->ncol * p.k);
  std::vector<double> dst((size_t)Q->ncol * p.k);
  int32_t keff = 0, warn = 0;
  const auto t0 = std::chrono::steady_clock::now();
  SD_TRY(sd_knn(ctx, Q, Y, &p, idx.data(), dst.data(), &keff, &warn));
  *sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return SD_OK;
}

int best_of(int reps, sd_context* ctx, sd_matrix* Q, sd_matrix* Y, int method, int prec,
            double* sec) {
  double best = 1e30;
  for (int r = 0; r < reps; ++r) {
    double t = 0;
    SD_TRY(time_knn(ctx, Q, Y, method, prec, &t));
    best = std::min(best, t);
  }
  *sec = best;
  return SD_OK;
}

bool plausible(double v, double lo, double hi) { return v == v && v >= lo && v <= hi; }

int calibrate(sd_context* ctx) {
  constexpr int64_t kRows = 16384, kRef = 32768;
  constexpr int kPer = 256;
  const SynthCsc Yh = synth_csc(kRows, kRef, kPer, 1);
  const SynthCsc Q1h = synth_csc(kRows, 256, kPer, 2);
  const SynthCsc Q2h = synth_csc(kRows, 8192, kPer, 3);
  const SynthCsc E1h = synth_csc(kRows, 8, kPer, 4);
  const SynthCsc E2h = synth_csc(kRows, 40, kPer, 5);
  const SynthCsc E0h = synth_csc(kRows, 1, 4 * kPer, 6);
  sd_matrix *Y = nullptr, *Q1 = nullptr, *Q2 = nullptr, *E1 = nullptr, *E2 = nullptr,
            *E0 = nullptr;
  auto release = [&]() {
    for (sd_matrix* m : {Y, Q1, Q2, E1, E2, E0})
      if (m) sd_matrix_free(m);
  };
  int rc = SD_OK;
  TuneRates t = ctx->tune;
  do {
    if ((rc = upload(ctx, kRows, Yh, &Y)) != SD_OK) break;
    if ((rc = upload(ctx, kRows, Q1h, &Q1)) != SD_OK) break;
    if ((rc = upload(ctx, kRows, Q2h, &Q2)) != SD_OK) break;
    if ((rc = upload(ctx, kRows, E1h, &E1)) != SD_OK) break;
    if ((rc = upload(ctx, kRows, E2h, &E2)) != SD_OK) break;
    if ((rc = upload(ctx, kRows, E0h, &E0)) != SD_OK) break;
    // densified path: two query batch sizes per kind; slope = rate, intercept = overhead
    const int methods[2] = {SD_COSINE, SD_BINARY};
    const int kinds[2] = {kDenseBf16Unit, kDenseF4Pattern};
    for (int m = 0; m < 2 && rc == SD_OK; ++m) {
      double w = 0, t1 = 0, t2 = 0;
      if ((rc = time_knn(ctx, Q1, Y, methods[m], SD_PREC_BF16, &w)) != SD_OK) break;  // warm-up
      if ((rc = best_of(3, ctx, Q1, Y, methods[m], SD_PREC_BF16, &t1)) != SD_OK) break;
      if ((rc = best_of(3, ctx, Q2, Y, methods[m], SD_PREC_BF16, &t2)) != SD_OK) break;
      const double kpad = (double)((kRows + dense_row_align(kinds[m]) - 1) /
                                   dense_row_align(kinds[m]) * dense_row_align(kinds[m]));
      const double w1 = 256.0 * kRef * 2.0 * kpad, w2 = 8192.0 * kRef * 2.0 * kpad;
      const double rate = (w2 - w1) / std::max(t2 - t1, 1e-6);
      if (plausible(rate, 1e12, 2e16)) {
        (m == 0 ? t.bf16_ops : t.fp4_ops) = rate;
        (m == 0 ? t.bf16_fixed_s : t.fp4_fixed_s) = std::max(t1 - w1 / rate, 1e-4);
      }
    }
    if (rc != SD_OK) break;
    // merge walk: two small query batches on the exact path
    double w = 0, e1 = 0, e2 = 0;
    if ((rc = time_knn(ctx, E1, Y, SD_COSINE, SD_PREC_EXACT, &w)) != SD_OK) break;
    if ((rc = best_of(3, ctx, E1, Y, SD_COSINE, SD_PREC_EXACT, &e1)) != SD_OK) break;
    if ((rc = best_of(3, ctx, E2, Y, SD_COSINE, SD_PREC_EXACT, &e2)) != SD_OK) break;
    const double s1 = 8.0 * kRef * 2.0 * kPer, s2 = 40.0 * kRef * 2.0 * kPer;
    const double steps = (s2 - s1) / std::max(e2 - e1, 1e-6);
    if (plausible(steps, 1e8, 1e14)) {
      t.merge_steps = steps;
      t.exact_fixed_s = std::max(e1 - s1 / steps, 5e-5);
    }
    // a lone query with 1024 stored rows: each pair walks 1280 steps in turn
    double e0 = 0;
    if ((rc = best_of(3, ctx, E0, Y, SD_COSINE, SD_PREC_EXACT, &e0)) != SD_OK) break;
    const double lat = (e0 - t.exact_fixed_s) / (4.0 * kPer + kPer);
    if (plausible(lat, 1e-9, 1e-4)) t.step_latency_s = lat;
    t.calibrated = 1;
  } while (false);
  release();
  if (rc == SD_OK) ctx->tune = t;
  return rc;
}

}  // namespace
}  // namespace sd

extern "C" int sd_autotune(sd_context* ctx, int32_t measure, sd_tune_info* out) {
  if (!ctx || !out) return sd::set_error(SD_EINVAL, "NULL argument");
  if (measure) {
    const sd_stats keep = ctx->stats;  // calibration runs are not the caller's searches
    int rc;
    try {
      sd::clear_error();
      rc = sd::calibrate(ctx);
    } catch (const std::bad_alloc&) {
      rc = sd::set_error(SD_ENOMEM, "host allocation failed");
    } catch (...) {
      rc = sd::set_error(SD_EDEVICE, "unexpected internal error");
    }
    ctx->stats = keep;
    if (rc != SD_OK) return rc;
  }
  const sd::TuneRates& t = ctx->tune;
  *out = sd_tune_info{t.bf16_ops,          t.fp4_ops,     t.merge_steps,
                      t.bf16_fixed_s * 1e3, t.fp4_fixed_s * 1e3, t.exact_fixed_s * 1e3,
                      t.calibrated,         0,             t.step_latency_s * 1e6};
  return SD_OK;
}

extern "C" int sd_context_set_tuning(sd_context* ctx, int32_t mode) {
  if (!ctx) return sd::set_error(SD_EINVAL, "NULL context");
  if (mode != 0 && mode != 1) return sd::set_error(SD_EINVAL, "tuning mode must be 0 or 1");
  ctx->sample_tuning = mode;
  return SD_OK;
}

extern "C" int sd_tune_samples(sd_tune_sample* out, int32_t max, int32_t* n, int32_t reset) {
  if (!n || (max > 0 && !out)) return sd::set_error(SD_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(sd::g_sample_mu);
  auto& t = sd::sample_table();
  int32_t i = 0;
  for (const auto& kv : t) {
    if (i < max) {
      sd_tune_sample& o = out[i];
      o.method = kv.first.method;
      o.layout = kv.first.kind;
      o.density_bucket = kv.first.density;
      o.row_bytes_log2 = kv.first.row_bytes;
      o.list_len = kv.first.L;
      o.symmetric = kv.first.tri;
      o.rule = kv.second.last_rule;
      o.last_run = kv.second.last_run;
      for (int c = 0; c < sd::kSampleCandidates; ++c) {
        const sd::SampleChoice sc = sd::sample_candidate(c, kv.second.last_L);
        o.stride[c] = sc.stride;
        o.rank[c] = sc.rank;
        o.ns_per_mpair[c] = kv.second.sec_per_pair[c] * 1e15;
        o.runs[c] = kv.second.runs[c];
      }
    }
    ++i;
  }
  *n = i;
  if (reset) t.clear();
  return SD_OK;
}
