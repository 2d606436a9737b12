// sd_internal.h -- shared declarations of libsd_hip.so (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sd.h"

namespace sd {

constexpr int32_t kNaInt = INT32_MIN;
constexpr uint64_t kNaRealBits = 0x7FF00000000007A2ULL;  // R's NA_real_

// ---------------------------------------------------------------------------
// error reporting (thread-local last error, as sd_last_error() documents)
// ---------------------------------------------------------------------------
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define SD_HIP(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return ::sd::set_error(e_ == hipErrorOutOfMemory ? SD_ENOMEM : SD_EDEVICE,      \
                             "HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                               \
  } while (0)

#define SD_TRY(expr)          \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != SD_OK) return rc_; \
  } while (0)

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
// A device allocation.  With `pool` set (the worker contexts of a server),
// memory comes from the device's stream-ordered pool on that stream
// (hipMallocAsync / hipFreeAsync), so growing a buffer never synchronises
// the device the way hipFree does; otherwise hipMalloc / hipFree.
// Device allocations made by DevBuf so far (a timed search that allocated
// is not a fair sample of its steady-state time).
extern std::atomic<int64_t> g_device_allocs;

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipStream_t pool = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes), pool(o.pool) {
    o.ptr = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr; bytes = o.bytes; pool = o.pool;
      o.ptr = nullptr; o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  int alloc(size_t n);            // exact size, old contents dropped
  int ensure(size_t n);           // grow-only
  void release();
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Device CSC view passed to kernels.
struct Csc {
  int64_t nrow = 0, ncol = 0;
  const int64_t* p = nullptr;
  const int32_t* i = nullptr;
  const double* x = nullptr;
};

}  // namespace sd

// ---------------------------------------------------------------------------
// opaque handles
// ---------------------------------------------------------------------------
namespace sd {
// Rates of the path chooser's cost model (sd_autotune measures them; the
// defaults were measured on an MI355X with the count-matrix workload).
struct TuneRates {
  double bf16_ops = 1.4e15;    // filtered pass, dense-equivalent ops/s (measured on MI355X)
  double fp4_ops = 5.2e15;
  double fp32_ops = 1.2e14;
  double merge_steps = 8.0e10; // exact slab merge-walk steps/s
  double step_latency_s = 0.4e-6;  // one pair's walk is sequential: time per step
  double bf16_fixed_s = 1.5e-3;  // per-search overheads
  double fp4_fixed_s = 0.5e-3;
  double fp32_fixed_s = 1.5e-3;
  double exact_fixed_s = 0.1e-3;
  int calibrated = 0;
};

// The filtered search's threshold sample (sd_tune.cpp): every stride-th
// reference tile of 128 columns, each query's rank-Lp key there.  The
// candidates trade pre-pass work (1/stride of the pairs) against the
// candidates the filtered pass appends (about stride x Lp per query).
constexpr int kSampleCandidates = 4;
struct SampleChoice {
  int stride = 32;
  int rank = 0;  // Lp
};
// Candidate c for the kept list length L: (32, L), (64, ~L/2), (128, ~L/4),
// (64, ~L/4 but at least 8).
SampleChoice sample_candidate(int c, int L);
// Searches alike in method, layout, reference density, row bytes, list
// length and symmetry share one tuning entry.
struct SampleKey {
  int method = 0, kind = 0, density = 0, row_bytes = 0, L = 0, tri = 0;
  bool operator<(const SampleKey& o) const;
};
// The candidate to run: the first allowed one not yet timed (when the
// search is large enough to time, *record is then set), else the fastest
// timed one; def when tuning is off or nothing is known.
int sample_plan(const SampleKey& key, const bool allowed[kSampleCandidates], int def, bool tuning,
                bool timeable, bool* record);
// Notes, for the listing, the rule's candidate and the one a search ran.
void sample_note(const SampleKey& key, int rule, int ran);
// Records a timed search (list length L): seconds per scored pair, from its
// pre-pass to its last re-rank or exact fallback.
void sample_record(const SampleKey& key, int cand, int L, double sec_per_pair);
}  // namespace sd

namespace sd {

// The candidate buffers of a second search, filled by the filtered pass of a
// cross search in the other direction: S(y, x) is S(x, y), so each value of
// the forward pass is also screened for the reverse search.
struct MirrorTarget {
  const double* thr = nullptr;   // per query of the reverse search
  int cap = 0;                   // buffer entries per query
  int32_t* cnt = nullptr;        // appended per query (zeroed by the pass)
  double* fkey = nullptr;        // [nqry][cap]
  int32_t* fidx = nullptr;
  const uint8_t* deg = nullptr;  // degenerate flags of its candidates (may be null)
  mutable bool filled = false;   // set by the pass that fed it
  // euclidean / pearson: the reverse search's dot error bound constants; the
  // pass feeds it only when they equal its own (its keys use them)
  float e0 = 0.f, e1 = 0.f;
  int nonneg = 0;
};

// A cross search run as one filtered pass (sd_knn_cross): the reverse
// search is probed first (its thresholds and buffers, no outputs), the
// forward search's filtered pass feeds those buffers, and the reverse search
// then resumes from them.  Any mismatch leaves the searches independent.
struct CrossMirror {
  int phase = 0;       // 0 off, 1 probe the reverse search, 2 forward, 3 resume the reverse
  bool ready = false;  // the probe set thresholds and buffers
  DevBuf thr, cnt, fkey, fidx;
  int L = 0, kind = -1, prec = 0, method = -1;
  int64_t kpad = 0, nqry = 0, nref = 0;
  MirrorTarget target;
};

}  // namespace sd

struct sd_context {
  int device = 0;
  sd::TuneRates tune;
  hipStream_t stream = nullptr;
  std::atomic<int> cancel{0};
  bool timing = false;
  sd_stats stats{};
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  int cu_count = 256;
  // scratch reused across calls (grow-only)
  sd::DevBuf scratch[28];
  // the SNN builder's temporaries, one slot each (grow-only)
  sd::DevBuf snn_scratch[24];
  sd::CrossMirror xm;
  std::vector<uint8_t> host_scratch;
  // pinned host memory that large results are copied into before the host
  // spreads them over the caller's arrays (grow-only; hipHostMalloc)
  void* host_stage = nullptr;
  size_t host_stage_bytes = 0;
  // device address of the allocation-free entry point's scratch array on
  // this context's device (resolved when the context is created)
  void* noalloc_scratch = nullptr;
  // the last self-search's neighbour indices (n x k column-major, as
  // returned), kept on the device for a following SNN call; knn_serial
  // counts the self-searches, 0 while none is kept
  sd::DevBuf knn_keep;
  int64_t knn_serial = 0, knn_calls = 0, keep_n = 0, keep_k = 0;
  // 1: the filtered search's threshold sample is tuned online per search
  // class (sd_tune.cpp); 0: the fixed rule
  int sample_tuning = 1;
  // device bytes one fused search may hold in candidate buffers and lists
  // before its queries are cut into blocks (0: half the free memory)
  int64_t buffer_budget = 0;
  // Capture of a search's launch sequence into a HIP graph (the server's
  // graph items): while set, a search synchronises nothing, reads nothing
  // back and takes no path that needs a host decision; it leaves the
  // filtered pass's overflow count and the certificate's failure count on
  // the device (cap_over, cap_fail; null when the search has none), which
  // the caller reads after each replay and answers with an ordinary search
  // when either is non-zero.
  bool capturing = false;
  const int32_t* cap_over = nullptr;
  const int32_t* cap_fail = nullptr;
};

struct sd_matrix {
  // the context it was loaded on; the matrix may outlive it, so only the
  // device ordinal is used afterwards
  sd_context* ctx = nullptr;
  int device = 0;
  // Held while derived layouts are built and until they are complete on the
  // building stream, so contexts on other streams (server workers) that
  // share the matrix never read a half-built layout.
  std::mutex prep_mu;
  int64_t nrow = 0, ncol = 0, nnz = 0;
  sd::DevBuf p, i, x;              // CSC (explicit zeros dropped)
  std::vector<int64_t> hp;         // host copy of colptr (planning)
  bool any_nonfinite = false;      // values contain NA/NaN/Inf
  bool any_negative = false;
  double abs_min = INFINITY, abs_max = 0.0;  // over finite stored values
  std::vector<uint8_t> hdeg;       // host copy of degenerate flags
  sd::DevBuf deg;                  // uint8 per column: all stored values == 0
  // cosine: unit-normalised values (nnz) and per-column max |v|
  bool unit_ready = false;
  sd::DevBuf unit, scale, unit_ssq;
  // row indices as uint16 (matrices of at most 65536 rows, built with the
  // unit values): the cosine re-rank streams half the index bytes
  sd::DevBuf rows16;
  // pearson / covariance: per-column mean and centred sum of squares, and
  // their FP32 forms for the slab epilogue (sqrt(nrow) mean, 1 / sqrt(css))
  bool centre_ready = false;
  sd::DevBuf mean, css, cmf, cif;
  // euclidean: per-column sum of squares (FP64, merge order irrelevant) and
  // the same rounded to FP32 for the slab epilogues
  bool sq_ready = false;
  sd::DevBuf sqn, sqf;
  // cosine re-rank margin tuned on this matrix as a reference (0 = default)
  int tuned_margin = 0;
  // euclidean / pearson on non-negative data start on plain BF16 values and
  // move to the split layout once a search on this reference had to fall
  // back to the exact path for more than 1% of its queries
  bool raw_split = false;
  // densified layouts, one per kind: [ncol][kpad] column-contiguous
  int64_t dense_kpad[7] = {0, 0, 0, 0, 0, 0, 0};
  sd::DevBuf dense[7];
  sd::Csc view() const {
    sd::Csc c;
    c.nrow = nrow; c.ncol = ncol;
    c.p = p.as<int64_t>(); c.i = i.as<int32_t>(); c.x = x.as<double>();
    return c;
  }
};

// whether prep_unit keeps the uint16 row indices: at most 65536 rows, and
// some column not fully stored (fully stored columns are re-ranked over their
// value arrays without row indices)
inline bool matrix_rows16(const sd_matrix* m) {
  return m->nrow <= 65536 && m->nnz > 0 && m->nnz < m->nrow * m->ncol;
}

struct sd_graph {
  int64_t n = 0, nnz = 0;
  int device = 0;
  sd::DevBuf colptr, rowind, val;
};

namespace sd {

// densified layout kinds.  kDenseBf16Split stores each value v as
// hi = bf16(v) and lo = bf16(v - hi) in three row segments [hi | hi | lo]; a
// query read with its last two segments swapped, [hi | lo | hi], gives the
// dot product hi.hi + hi.lo + lo.hi in one MFMA pass over 3 kpad.
// kDenseBf16Raw holds bf16(v) only: for non-negative data the keys' error
// bound is relative to each pair's own dot product, which plain BF16 meets.
enum DenseKind { kDenseF4Pattern = 0, kDenseBf16Unit = 1, kDenseBf16Split = 2, kDenseF32Unit = 3,
                 kDenseF32Raw = 4, kDenseF32Pattern = 5, kDenseBf16Raw = 6, kDenseKinds = 7 };
static_assert(sizeof(sd_matrix::dense) / sizeof(DevBuf) == kDenseKinds, "one layout slot per kind");

// Serve every device buffer of ctx (and of matrices loaded on it with
// matrix_load into buffers given the same pool) from the device's
// stream-ordered pool on ctx's stream, which keeps freed memory for reuse.
int context_use_pool(sd_context* ctx);
void matrix_use_pool(sd_matrix* m, hipStream_t stream);

// Validate a host CSC matrix and copy it into m (device buffers grown only
// when too small; derived layouts of earlier contents are discarded).
int matrix_load(sd_context* ctx, sd_matrix* m, int64_t nrow, int64_t ncol, const int64_t* colptr,
                const int32_t* rowind, const double* values);

// hipStreamSynchronize of ctx's stream, skipped while a search is captured
int sync_stream(sd_context* ctx);
// n int32 zeros at p on ctx's stream by a kernel (a captured graph gets no
// memset nodes)
int zero_i32(sd_context* ctx, int32_t* p, int64_t n);
// Forget m's derived layouts (unit values, statistics, densified columns)
// without freeing their buffers, so the next search builds them again in
// place.
void matrix_forget_derived(sd_matrix* m);
// One search of X against Y (null: self-search) whose results stay on the
// device: nq x k_eff column-major at *d_idx / *d_val (ctx's output buffers,
// valid until ctx's next search).  With ctx->capturing the launches go into
// the stream's capture; a search that would need a host decision or an
// allocation returns SD_EUNSUPPORTED.
int knn_device(sd_context* ctx, sd_matrix* X, sd_matrix* Y, const sd_knn_params* p,
               int32_t* k_eff, int32_t* warn, const int32_t** d_idx, const double** d_val);

// ---- prep (k_prep.hip) ------------------------------------------------------
int prep_flags(sd_context* ctx, sd_matrix* m);                 // deg, nonfinite, negative
int prep_unit(sd_context* ctx, sd_matrix* m);                  // cosine unit values
int prep_centre(sd_context* ctx, sd_matrix* m);                // pearson/covariance stats
int prep_sq(sd_context* ctx, sd_matrix* m);                    // euclidean squared norms
int prep_dense(sd_context* ctx, sd_matrix* m, int kind, int64_t kpad);

// ---- exact FP64 path (k_exact.hip) -------------------------------------------
struct ExactArgs {
  int method = 0, dist = 1;
  Csc ref, qry;                 // full matrices
  const double* ref_unit = nullptr; const double* ref_scale = nullptr;
  const double* qry_unit = nullptr; const double* qry_scale = nullptr;
  const double* ref_mean = nullptr; const double* ref_css = nullptr;
  const double* qry_mean = nullptr; const double* qry_css = nullptr;
  const uint8_t* ref_deg = nullptr;   // nullptr: no masking
  const uint16_t* ref_rows16 = nullptr;  // ref.i as uint16, or nullptr
  int self_mode = 0;            // query j is reference column j + self_off
  int64_t self_off = 0;
  int include_self = 0;
};
// D[i + b*ld] for refs i in [0, ref.ncol) and queries qcols[b] (b < nb).
int exact_slab(sd_context* ctx, const ExactArgs& a, const int32_t* d_qcols, int64_t nb, double* D,
               int64_t ld, int mask_self_and_deg);

// Top-k of dense device columns D (nrow x ncol, leading dim ld): keys ranked
// by (value, row) with the method's direction; rows listed in d_skip (per
// column, 0-based, -1 none) and non-finite values are skipped.  Writes
// 1-based idx / values into out (ncol x k, leading dim out_ld) at column
// offsets given by d_outcol (nullptr: identity).
int topk_dense(sd_context* ctx, const double* D, int64_t nrow, int64_t ncol, int64_t ld, int k,
               int decreasing, const int32_t* d_skip, int32_t* out_idx, double* out_val,
               int64_t out_ld, const int32_t* d_outcol);

// ---- fused MFMA path (k_slab.hip) ------------------------------------------
struct FusedArgs {
  int method = 0;               // SD_BINARY / SD_COSINE / SD_EUCLIDEAN / SD_PEARSON
  int prec = SD_PREC_BF16;
  int kind = kDenseBf16Unit;    // densified layout of ref_dense / qry_dense
  int64_t nref = 0, nqry = 0, kpad = 0;
  const void* ref_dense = nullptr;   // chunk-major layouts (dense_offset) ...
  const void* qry_dense = nullptr;
  int64_t ref_ld = 0, qry_ld = 0;   // ... of ref_ld / qry_ld columns per chunk block
  const int32_t* qlist = nullptr;   // optional: queries are these columns (positions 0..nqry)
  const int64_t* ref_p = nullptr;   // column pointers (nnz per column)
  const int64_t* qry_p = nullptr;
  // per-column FP32 aux: euclidean ||x||^2; pearson sqrt(nrow) mean (aux)
  // and 1 / sqrt(centred sum of squares) (aux2, 0 for a constant column)
  const float* ref_aux = nullptr;
  const float* qry_aux = nullptr;
  const float* ref_aux2 = nullptr;
  const float* qry_aux2 = nullptr;
  // euclidean / pearson keys: the raw dot product's error bound is
  // (dot_e0 + dot_e1 nnz(q)) times sum |q_i r_i| (euc_key, pear_key in
  // k_slab.hip); dot_nonneg: every value of both matrices is >= 0, so that
  // sum is the dot product itself
  float dot_e0 = 0.f, dot_e1 = 0.f;
  int dot_nonneg = 0;
  const uint8_t* ref_deg = nullptr;
  int self_mode = 0, include_self = 0;
  int64_t self_off = 0;             // self mode: query j is reference column j + self_off
  int ref_density_bucket = 0;       // round(-2 log2(stored fraction of the reference))
  int L = 0;                        // list length kept per query
  int splits = 1;
  int tile_stride = 1;              // list mode: only every tile_stride-th reference tile
  const MirrorTarget* mirror = nullptr;  // filtered pass: also feed the reverse search
};
// Candidate lists: per (split, query) L entries of (key double, idx int32),
// key larger = better, sorted best first; unfilled slots idx = -1.
int fused_slab_topl(sd_context* ctx, const FusedArgs& a, double* cand_key, int32_t* cand_idx);

struct RerankArgs {
  ExactArgs ex;
  const double* cand_key;      // merged candidate lists, nqry x L
  const int32_t* cand_idx;
  int64_t nqry;
  int L, k, decreasing;
  int nonneg;                  // both matrices non-negative
  double in_round;             // relative rounding of each slab input
  double acc_round;            // unit roundoff of the accumulator
  const uint8_t* qdeg;
  int32_t* out_idx;
  double* out_val;
  int32_t* fail_list;          // queries the certificate could not clear
  int32_t* fail_count;
  const int32_t* qlist = nullptr;  // optional: position -> query column
  int64_t out_ld = 0;          // leading dimension of out_* (all query columns)
  int map_words = 0;           // LDS row bitmap words (0: merge walk)
};
int rerank_map_words(int64_t nrow);
// Densified columns are padded with zero rows so that every slab kernel's
// K loop runs whole groups of chunks: a padded column is a multiple of 256
// bytes (the slab kernels stream 128-byte chunks in pairs) and of 256 rows.
// The FP4 pattern layout packs two rows per byte.
//
// The densified matrix is stored chunk-major: K byte b of column j of a
// matrix with ld columns sits at ((b / 128) ld + j) 128 + b % 128, so the
// 128-byte chunk c of a tile of consecutive columns is one contiguous block
// (32 KB for 256 columns).  A slab kernel streams a tile's chunks in order:
// each request then reads whole DRAM pages instead of 128 bytes at a stride
// of the padded column.
constexpr int64_t kDenseChunk = 128;
__host__ __device__ inline int64_t dense_offset(int64_t col, int64_t kbyte, int64_t ld) {
  return ((kbyte >> 7) * ld + col) * kDenseChunk + (kbyte & (kDenseChunk - 1));
}
inline int64_t dense_row_align(int kind) { return kind == kDenseF4Pattern ? 512 : 256; }
inline int64_t dense_row_bytes(int kind, int64_t kpad) {
  return kind == kDenseF4Pattern                             ? kpad / 2
         : (kind == kDenseBf16Unit || kind == kDenseBf16Raw) ? kpad * 2
         : kind == kDenseBf16Split                           ? kpad * 6
                                                             : kpad * 4;
}
// Bytes of one segment whose query copy is read swapped with the next one
// (the split layout's hi | lo), 0 for layouts read as they are.
inline int64_t dense_qseg_bytes(int kind, int64_t kpad) {
  return kind == kDenseBf16Split ? kpad * 2 : 0;
}
// Cost-model choice between the densified path and the FP64 merge walk for
// a search of nq query columns against nref reference columns (symmetric: a
// self-search whose filtered pass scores each pair once).
bool tune_prefer_exact(const sd_context* ctx, int kind, int64_t nq, int64_t nref, int64_t kpad,
                       double nnz_q, double nnz_ref, bool symmetric);
int fused_list_len(int want);
// Whether a slab tile of 256 columns of row_bytes each stays within the
// 32-bit buffer offsets of the slab kernels.
bool fused_fits(int64_t row_bytes);
// Filter mode: every (query tile, reference tile) once; candidates with key
// >= thr[query] are appended to fkey/fidx ([nqry][cap]), counted in cnt.
int fused_filter_pass(sd_context* ctx, const FusedArgs& a, const double* thr, int cap,
                      int32_t* cnt, double* fkey, int32_t* fidx);
// Per-query top-L of the appended candidates.  Queries with more than cap
// candidates (buffer overflow) or fewer than min_count (the threshold cannot
// vouch for complete lists) are counted in over_count and, when over_list is
// given, listed there for a full pass.  With qlist, list position b takes
// query qlist[b]'s buffer (outputs by position) and over_list records the
// positions.
int select_lists(sd_context* ctx, const double* fkey, const int32_t* fidx, const int32_t* cnt,
                 int cap, int64_t nqry, int L, double* out_key, int32_t* out_idx,
                 int32_t* over_list, int32_t* over_count, int min_count,
                 const int32_t* qlist = nullptr);
int list_thresholds(sd_context* ctx, const double* key, const int32_t* idx, int64_t nqry, int L,
                    double* thr);
int scatter_lists(sd_context* ctx, const double* key, const int32_t* idx, const int32_t* rows,
                  int64_t n, int L, double* out_key, int32_t* out_idx);
// out[j] = src[pos[j]] for j < n (device arrays).
int gather_i32(sd_context* ctx, const int32_t* src, const int32_t* pos, int64_t n, int32_t* out);
int fused_effective_splits(const FusedArgs& a);
int merge_lists(sd_context* ctx, const double* key, const int32_t* idx, int nlists, int64_t nqry,
                int L, double* out_key, int32_t* out_idx);
// NA rows (index and distance) for the degenerate query columns of a
// column-major nq x k result.
int fill_degenerate(sd_context* ctx, int32_t* idx, double* val, const uint8_t* deg, int64_t nq,
                    int k);
// Jaccard results (nqry x k, leading dimension out_ld) from exact lists.
int finish_exact_lists(sd_context* ctx, const double* key, const int32_t* idx, int64_t nqry, int L,
                       int k, int dist, const uint8_t* qdeg, int32_t* out_idx, double* out_val,
                       int64_t out_ld);
int rerank_cosine(sd_context* ctx, const RerankArgs& R);
// Euclidean: exact FP64 distances of the candidates in the reference's
// arithmetic; the certificate compares the k-th squared distance with the
// lower bound that the list's last key stands for.
int rerank_euclid(sd_context* ctx, const RerankArgs& R);
// Pearson: exact FP64 correlations (the library's centred definition); the
// certificate compares the k-th correlation with the upper bound that the
// list's last key stands for.
int rerank_pearson(sd_context* ctx, const RerankArgs& R);
int topk_rows(sd_context* ctx, const double* D, int64_t nrow, int64_t ncol, int k, int decreasing,
              int32_t* best_idx, double* best_val, int64_t col_offset, const int32_t* self_col);

// ---- SNN (k_snn.hip) ---------------------------------------------------------
// nids = 0: neighbour ids are 1-based node indices (the reference's sparseSNN
// input).  nids > 0: they are 1-based columns of another matrix with nids
// columns (a query batch's neighbours in a reference); only edges = 1
// ("shared") applies and a node is never its own neighbour.
int snn_build(sd_context* ctx, const int32_t* d_idx, int n, int k, int include_self, double prune,
              int symmetrise, int edges, sd_graph* g, double* d_w_out, int nids = 0);

// One-workgroup-per-item launches are cut into grids of at most this many
// workgroups: a dispatch packet counts its work-items in 32 bits, so a 1-D
// grid of 256-thread workgroups must stay below 2^24 of them.
constexpr int64_t kMaxItemsPerLaunch = (int64_t)1 << 22;
// Grid of an elementwise kernel that strides over n items (capped; the
// kernels loop).
inline unsigned stride_grid(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > (1 << 20) ? (1 << 20) : b));
}

// Device address of the allocation-free SNN entry point's static scratch
// array on `device` (the current device), resolved once per device.
int noalloc_scratch_setup(int device, void** out);

// ---- timing helpers ------------------------------------------------------------
struct KernelTimer {
  sd_context* ctx;
  double* acc;
  int64_t* count;
  bool on;
  KernelTimer(sd_context* c, double* a, int64_t* n);
  ~KernelTimer();
};

}  // namespace sd
