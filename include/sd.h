/*
 * sd.h -- C ABI of libsd_hip.so, the MI355X (gfx950) implementation of
 * sparseDist's blocked exact-kNN and SNN path.
 *
 * The reference binds its native code through Rcpp-generated .Call routines
 * (src/RcppExports.cpp:306-333).  Each entry point below names the reference
 * interface it replaces.  Only plain C types cross this boundary: no R, Rcpp,
 * Armadillo or torch types, no exceptions.  Every function returns an int
 * status (SD_OK = 0); the message of the last failure on the calling thread is
 * available from sd_last_error().
 *
 * Conventions shared by every entry point:
 *   - Sparse input is CSC (dgCMatrix layout): colptr int64[ncol+1], rowind
 *     int32[nnz] strictly ascending within a column, values double[nnz].
 *     Explicit zeros are dropped at upload (the reference's dgCMatrix ->
 *     arma::sp_mat conversion stores none), so Jaccard counts structural
 *     non-zeros; "degenerate" (all-zero) columns are detected on the values
 *     as R/sparseKNN.R:214-217 does.
 *   - kNN results are FNN-style: idx int32 and dist double, nq x k
 *     column-major (entry (q, t) at t*nq + q), 1-based indices, best first.
 *     Empty slots hold R's NA_integer_ (INT32_MIN) in idx and R's NA_real_
 *     (bit pattern 0x7FF00000000007A2) in dist, as src/topK.cpp:96-101, 140-145.
 *   - Ties are broken by the lower column index across every tile, slab and
 *     launch (src/topK.cpp:42-49).  Results never depend on block_size, tile
 *     shape or batching.
 *   - Output pointers are host memory unless a parameter says otherwise.
 *   - One context may be used from one host thread at a time; sd_cancel() is
 *     the exception and may be called from any thread.
 */
#ifndef SD_H
#define SD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
enum {
  SD_OK = 0,
  SD_EINVAL = 1,       /* invalid argument (text as the reference's R/C++ errors) */
  SD_ENONFINITE = 2,   /* "Cosine requires finite input (found NA, NaN or Inf)." */
  SD_ENOMEM = 3,       /* device or host allocation failed */
  SD_ENOCAND = 4,      /* "No candidate neighbours available." */
  SD_ECAP = 5,         /* caller buffer too small */
  SD_ECANCELLED = 6,   /* "Computation interrupted by the user." */
  SD_EDEVICE = 7,      /* HIP runtime error */
  SD_EUNSUPPORTED = 8, /* method or option not provided by this library */
  SD_EBUDGET = 9       /* serving: request larger than the memory budget */
};

/* ---- methods (R/sparseKNN.R:156-160; "jaccard" is an alias of binary) ----- */
enum {
  SD_BINARY = 0,
  SD_COSINE = 1,
  SD_EUCLIDEAN = 2,
  SD_MANHATTAN = 3,
  SD_PEARSON = 4,
  SD_COVARIANCE = 5,
  SD_JS = 6          /* Jensen-Shannon distance: finite, non-negative input */
};

/* ---- slab arithmetic ------------------------------------------------------ */
enum {
  SD_PREC_AUTO = 0,  /* BF16 MFMA + FP64 re-rank where it applies, else exact */
  SD_PREC_EXACT = 1, /* FP64 merge walk for every pair (the reference's order) */
  SD_PREC_FP32 = 2,  /* plain FP32 dense slab + FP64 re-rank (baseline path) */
  SD_PREC_BF16 = 3   /* BF16 MFMA, FP32 accumulate, FP64 re-rank + certificate */
};

typedef struct sd_context sd_context;
typedef struct sd_matrix sd_matrix;
typedef struct sd_graph sd_graph;

const char* sd_last_error(void);

/* Library version string (also records the build's offload target). */
const char* sd_version(void);

/* ---- context -------------------------------------------------------------- */
/* Owns the HIP stream, scratch memory and statistics.  device = HIP ordinal. */
int sd_context_create(int device, sd_context** out);
int sd_context_destroy(sd_context* ctx);

/* Request cancellation of the running call (reference Progress::check_abort /
 * stop_if_aborted, src/sparseDist.cpp:177-179).  The call polls the request
 * between kernels, returns SD_ECANCELLED and writes no partial output.  A
 * request made while no call runs applies to the next call on the context
 * (a caller's interrupt poll may run before its worker thread enters the
 * call); every call clears the request when it returns.  Callable from any
 * thread. */
int sd_cancel(sd_context* ctx);

/* Kernel statistics since the last reset, measured with HIP events on the
 * context's stream while timing is on: device time in ms and launch count of
 * the fused slab kernel in list mode (threshold pre-pass, small searches and
 * the wider second pass), of its filtered full pass over every reference
 * tile, of the FP64 re-rank and of the SNN builder; how many query columns
 * the certificate cleared, how many needed the wider second pass, how many
 * went to the exact slab and how many overflowed the candidate buffer. */
typedef struct {
  double slab_ms;
  int64_t slab_launches;
  double rerank_ms;
  int64_t rerank_launches;
  double snn_ms;
  int64_t snn_launches;
  int64_t fallback_queries;   /* sent to the exact FP64 slab */
  int64_t certified_queries;  /* cleared by the certificate (first or second pass) */
  int64_t escalated_queries;  /* needed the second, wider candidate pass */
  int64_t overflow_queries;   /* candidate buffer overflowed: full per-query pass */
  double filter_ms;           /* filtered full pass (not included in slab_ms) */
  int64_t filter_launches;
  int64_t fused_calls;        /* searches run on the densified MFMA path */
  int64_t exact_calls;        /* searches run on the FP64 merge walk */
  int64_t last_margin;        /* re-rank margin m of the last fused search */
  int64_t slab_products;      /* MFMA products per row of the last fused search's layout:
                                 3 for split BF16 (hi.hi + hi.lo + lo.hi), else 1 */
  int64_t filter_pairs;       /* (query, reference) pairs scored by filtered passes: a
                                 self-search over all columns at once scores each
                                 unordered pair once */
  int64_t buffer_escalations; /* searches whose wider second lists came from the filtered
                                 pass's candidate buffers instead of a second slab pass */
  int64_t escalation_list_queries; /* of those, queries whose buffer could not vouch for the
                                      wider list (too few entries, or overflowed) and took a
                                      list-mode pass of their own */
  int64_t query_blocks;       /* blocks of query columns the fused searches ran in (one per
                                 search unless its buffers exceeded the buffer budget) */
  int64_t snn_device_inputs;  /* sd_snn calls that read their indices from device memory */
  int64_t sample_stride;      /* threshold sample of the last filtered search: tile stride */
  int64_t sample_rank;        /* ... and rank Lp */
} sd_stats;
int sd_context_stats(sd_context* ctx, sd_stats* out, int reset);
/* Enable/disable event timing of kernels (off by default). */
int sd_context_set_timing(sd_context* ctx, int on);
/* Device bytes one fused search may hold in candidate buffers and lists
 * (about 50 KB per query column at most).  A search whose query columns need
 * more runs them in blocks, each a whole search of its own with identical
 * results; 0 (the default) means half the device memory free at the call.
 * No counterpart in the reference (which keeps one dense block_size x n_ref
 * slab per pass, R/sparseKNN.R:68 block_size). */
int sd_context_set_buffer_budget(sd_context* ctx, int64_t bytes);
/* The context's hipStream_t, for callers that order their own work. */
void* sd_context_stream(sd_context* ctx);

/* ---- autotuning ------------------------------------------------------------ */
/* With SD_PREC_AUTO each search picks its path from a cost model: the
 * densified MFMA path costs a fixed overhead plus 2 r n_ref per query column
 * (whole query tiles of 256; half that for a self-search, which scores each
 * pair once) at the filtered pass's rate; the FP64 merge
 * walk costs a smaller overhead plus n_ref (nnz_q + nnz_ref) merge steps per
 * query column.  Both paths give identical results, so the choice never
 * changes an answer.  The re-rank margin m (cosine, euclidean, pearson)
 * adapts per reference matrix.  When the wider second lists come from the
 * filtered pass's candidate buffers, a search whose certificate sends more
 * than 15% of its queries to the second re-rank doubles m for the next search
 * on that reference (up to 48), and one that sends fewer than 1% lowers it
 * by 4 (down to 8).  When they need a second slab pass, more than 1% doubles
 * m and a search with none lowers it by 8 (down to 16).  The reference
 * has no counterpart; its bench chooses between dense and sparse inputs by a
 * density sweep (bench/R/20-design.R:115-160).
 * sd_autotune measures the model's rates on this device (about a second of
 * work on synthetic inputs) and keeps them in the context; until it is
 * called, measured MI355X defaults are used. */
typedef struct {
  double bf16_ops;   /* filtered-pass dense-equivalent ops/s, BF16 values */
  double fp4_ops;    /* same, FP4 0/1 patterns */
  double merge_steps; /* FP64 merge-walk steps/s of the exact slab */
  double bf16_fixed_ms;   /* per-search overhead of the densified cosine path */
  double fp4_fixed_ms;    /* per-search overhead of the densified Jaccard path */
  double exact_fixed_ms;  /* per-search overhead of the merge walk */
  int32_t calibrated;     /* 1 once sd_autotune has measured them */
  int32_t reserved;
  double step_latency_us; /* merge walk: time per sequential step of one pair */
} sd_tune_info;
int sd_autotune(sd_context* ctx, int32_t measure, sd_tune_info* out);

/* The filtered search's threshold sample is tuned online.  Large searches
 * run in two passes: a list-mode pre-pass over every stride-th reference
 * tile gives each query its rank-Lp key there as a threshold, and the
 * filtered pass over all tiles keeps the candidates at or above it.  A
 * sparser sample costs the pre-pass less and leaves the filtered pass more
 * candidates (about stride x Lp per query); the balance depends on the row
 * length, the density, the list length and the symmetry of the pass.  Each
 * class of searches (method, layout, density bucket, row bytes, list length
 * class, symmetric or not) times four candidates -- (stride, Lp) = (32, L),
 * (64, ~L/2), (128, ~L/4) and (64, ~L/4 but >= 8) -- on its first searches
 * of at least 2^28 scored pairs, one candidate each, and runs the fastest
 * afterwards.  Results do not depend on the choice.  The table is
 * process-wide.  No counterpart in the reference, whose design bench sweeps
 * density to place its dense/sparse crossover (bench/R/20-design.R:115-160).
 * sd_context_set_tuning: 1 (default) tunes; 0 uses the fixed rule (every
 * 64th tile at ~L/4 for rows of at most 2 KB, else 128th / 64th / 32nd by the
 * reference's tile count). */
int sd_context_set_tuning(sd_context* ctx, int32_t mode);
typedef struct {
  int32_t method, layout;     /* SD_* method; densified layout code */
  int32_t density_bucket;     /* round(-2 log2(stored fraction of the reference)) */
  int32_t row_bytes_log2;     /* ceil(log2(bytes per densified column)) */
  int32_t list_len;           /* list length class: 16 (L <= 16) or 64 (longer lists) */
  int32_t symmetric;          /* 1: a self-search's symmetric pass */
  int32_t rule;               /* candidate the fixed rule chose for the class's last search */
  int32_t last_run;           /* candidate that search ran (-1: none noted yet) */
  int32_t stride[4];          /* the candidates: sampled tile stride ... */
  int32_t rank[4];            /* ... and rank Lp (for the class's last timed list length) */
  double ns_per_mpair[4];     /* measured ns per 10^6 scored pairs (0: not yet timed) */
  int64_t runs[4];            /* timed searches per candidate */
} sd_tune_sample;
/* Copies up to max entries of the sample table; *n = the number of entries.
 * reset != 0 clears the table afterwards. */
int sd_tune_samples(sd_tune_sample* out, int32_t max, int32_t* n, int32_t reset);

/* ---- matrices ------------------------------------------------------------- */
/* Copy a CSC matrix to device memory.  Replaces the per-call dgCMatrix ->
 * arma::sp_mat conversion at the reference's .Call boundary
 * (src/RcppExports.cpp:124-137, 229-242; R/sparseDist.R:198-202): the matrix
 * stays resident and its derived layouts (norms, densified tiles, and for
 * cosine on matrices of at most 65536 rows a uint16 copy of the row indices,
 * 2 bytes per stored entry) are cached with it. */
int sd_matrix_upload_csc(sd_context* ctx, int64_t nrow, int64_t ncol, const int64_t* colptr,
                         const int32_t* rowind, const double* values, sd_matrix** out);
int sd_matrix_free(sd_matrix* m);
/* Drop cached derived layouts (the CSC itself stays resident). */
int sd_matrix_drop_cache(sd_matrix* m);

/* ---- kNN ------------------------------------------------------------------ */
typedef struct {
  int32_t k;               /* neighbours requested (>= 1) */
  int32_t method;          /* SD_BINARY ... SD_JS */
  int32_t dist;            /* 1 = distance form, 0 = similarity form */
  int32_t include_self;    /* self-search only (ignored when Y != NULL) */
  int32_t drop_degenerate; /* mask all-zero columns on both sides */
  int32_t precision;       /* SD_PREC_* */
  int32_t margin;          /* extra candidates re-ranked in FP64 (0 = default) */
  int32_t reserved;
  int64_t block_size;      /* query columns per slab: a memory hint only */
} sd_knn_params;

/* Blocked exact kNN of the columns of X among the columns of Y (Y == NULL:
 * self-search among X).  Replaces R/sparseKNN.R:110-313 after argument
 * validation, i.e. the block loop over sparseDist() (fastJacc2
 * src/sparseDist.cpp:559-617, fastCosine2 967-1000, fastEuclidean2 833-878,
 * fastManhattan2 734-779, fastCorr2 403-441, fastCov2 481-510, fastJS2
 * 1074-1115) and topKBlock
 * (src/topK.cpp:67-149), fused so the slab never leaves the device.
 * idx/dist have room for X.ncol * params.k entries; the result occupies the
 * first X.ncol * (*k_eff) entries.  *warn = 1 when k was clamped to the
 * number of available candidates (the caller raises the reference's
 * "k (K) exceeds the A available neighbours; returning A." warning). */
int sd_knn(sd_context* ctx, const sd_matrix* X, const sd_matrix* Y, const sd_knn_params* params,
           int32_t* idx, double* dist, int32_t* k_eff, int32_t* warn);

/* Both directions (R/sparseKNNCross.R:44-199): forward = kNN of X among Y,
 * reverse = kNN of Y among X, each bit-identical to the corresponding sd_knn
 * call (the reference's equivalence contract, R/sparseKNNCross.R:33-40).
 * Replaces the slab + topKBlock + topKRowsAccum (src/topKRows.cpp:104-248)
 * loop.  When both sides are large enough for a filtered pass (at least
 * 16384 columns each) and both directions take the same layout, one pass
 * over the X x Y tiles serves both (the slab is the same, transposed).
 * warn bit 0: forward k clamped; bit 1: reverse k clamped. */
int sd_knn_cross(sd_context* ctx, const sd_matrix* X, const sd_matrix* Y,
                 const sd_knn_params* params, int32_t* fwd_idx, double* fwd_dist, int32_t* k_fwd,
                 int32_t* rev_idx, double* rev_dist, int32_t* k_rev, int32_t* warn);

/* ---- one-to-one replacements of the reference's native helpers ------------ */
/* Dense slab D (X.ncol x Y.ncol, column-major, host memory) in FP64 with the
 * reference's operation order.  Replaces fastJacc2 / fastCosine2 /
 * fastEuclidean2 / fastManhattan2 / fastCorr2 / fastCov2 / fastJS2
 * (src/RcppExports.cpp:80-93, 110-137, 199-242; R/RcppExports.R:24-66) with
 * m = X and m2 = Y (fastJS2 takes dense matrices; their zero entries are the
 * ones this CSC does not store). */
int sd_slab(sd_context* ctx, int32_t method, const sd_matrix* X, const sd_matrix* Y, int32_t dist,
            double* D);

/* Column-wise top-k of a dense host block D (nrow x ncol).  Replaces
 * topKBlock (src/topK.cpp:67-149, src/RcppExports.cpp:273-286).  self_row:
 * per column, the 1-based row to exclude or 0.  idx/val are ncol x k. */
int sd_topk_block(sd_context* ctx, const double* D, int64_t nrow, int64_t ncol, int32_t k,
                  int32_t decreasing, const int32_t* self_row, int32_t* idx, double* val);

/* Fold a dense host block into a running per-row top-k held in host memory
 * (best_idx/best_val, nrow x k, updated in place).  Replaces topKRowsAccum
 * (src/topKRows.cpp:104-248, src/RcppExports.cpp:288-304). */
int sd_topk_rows_accum(sd_context* ctx, const double* D, int64_t nrow, int64_t ncol, int32_t k,
                       int32_t decreasing, int32_t* best_idx, double* best_val,
                       int64_t col_offset, const int32_t* self_col);

/* ---- SNN ------------------------------------------------------------------ */
/* Jaccard weights over the kNN edge set: idx n x k (1-based, NA_integer_
 * padding) -> w n x k (NA_real_ where undefined).  Replaces snnJaccard
 * (src/snn.cpp:198-250, src/RcppExports.cpp:35-47). */
int sd_snn_jaccard(sd_context* ctx, const int32_t* idx, int32_t n, int32_t k,
                   int32_t include_self, double* w);

/* sd_snn_jaccard without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * d_idx (n x k int32, 1-based, NA_integer_ padding) and d_w (n x k double)
 * are DEVICE pointers; the weights are those of sd_snn_jaccard bit for bit.
 * The neighbour sets live in a static device array of 64 MiB per GPU, held
 * until the process ends: n * (ks + 1) * 4 + 768 bytes must fit in it, with
 * ks = k + 1 rounded up to a multiple of 4 (n <= 671,000 at k = 20);
 * larger inputs are rejected with SD_EINVAL before any work.  All calls on
 * one GPU share that array, so calls that overlap on one GPU race.  Runs on
 * ctx's stream and returns when the weights are written. */
int sd_snn_jaccard_noalloc(sd_context* ctx, const int32_t* d_idx, int32_t n, int32_t k,
                           int32_t include_self, double* d_w);

/* The SNN graph as an n x n sparse matrix on the device.  edges: 0 = "knn"
 * (weights over the kNN edge set, strict w > prune, duplicates collapsed,
 * optional union symmetrisation; R/sparseSNN.R:221-250), 1 = "shared"
 * (every pair sharing a neighbour, pruned in the kernel; replaces
 * snnJaccardShared src/snn.cpp:299-447 and R/sparseSNN.R:206-218).
 * idx_on_device: idx is a device pointer (e.g. an sd_knn result kept on the
 * GPU) rather than host memory. */
int sd_snn(sd_context* ctx, const int32_t* idx, int32_t n, int32_t k, int32_t include_self,
           double prune, int32_t symmetrise, int32_t edges, int32_t idx_on_device,
           sd_graph** out);
/* The neighbour indices of the context's last sd_knn self-search (Y NULL),
 * kept in device memory exactly as that call returned them (n x k
 * column-major int32, 1-based, NA_INTEGER for missing neighbours) until the
 * next self-search on the context replaces them: pass *d_idx to sd_snn with
 * idx_on_device = 1 instead of uploading the host copy again.  *serial
 * numbers the self-searches of the context; a caller that records it right
 * after its sd_knn call can tell whether the kept indices are still its
 * result.  *d_idx is NULL and *serial 0 while nothing is kept.  No
 * counterpart in the reference, where sparseSNN(nn) reads nn$idx
 * (R/sparseSNN.R:124-166). */
int sd_knn_kept(sd_context* ctx, const int32_t** d_idx, int32_t* n, int32_t* k, int64_t* serial);
/* Size of the graph: n and stored entries. */
int sd_graph_info(const sd_graph* g, int64_t* n, int64_t* nnz);
/* Copy the graph to host memory in dgCMatrix layout: colptr int64[n+1],
 * rowind int32[nnz] (0-based, ascending per column), val double[nnz]. */
int sd_graph_copy(const sd_graph* g, int64_t* colptr, int32_t* rowind, double* val);
int sd_graph_free(sd_graph* g);

/* ---- serving ---------------------------------------------------------------
 * A request loop over one reference matrix Y kept resident in device memory
 * with its norms and densified layouts.  The reference has no counterpart:
 * its closest analogue is the block loop of R/sparseKNN.R:248-294, which
 * re-converts the whole reference for every block.  Each request is one
 * sd_knn(X_request, Y) call (same parameters, same results bit for bit),
 * optionally followed by the shared-neighbour graph of its query columns.
 *
 *  - Admission: a request reserves an estimate of the device memory its work
 *    needs (one chunk's densified columns and candidate buffers, its results
 *    and SNN workspace).  Requests are admitted in arrival order while the
 *    reservations fit in budget_bytes; one that does not fit yet lets later
 *    requests that fit pass it until it has waited 2 s or been passed 64
 *    times, after which it is admitted next.  Requests of at most 1000
 *    columns that the budget cannot take yet (or that wait behind such a
 *    request) may use a lane of budget_bytes / 16 beyond the budget, so a
 *    large request holding the queue delays only other large ones.  A
 *    request that cannot fit even alone, or that arrives while max_pending
 *    requests wait, is refused with SD_EBUDGET.
 *  - Memory: each worker reloads its query columns into buffers it keeps,
 *    and grows its scratch from the device's stream-ordered pool, so items
 *    do not allocate or free device memory once warm.
 *  - Continuous batching: `slots` workers, each with its own context and
 *    stream, take work items of at most chunk_cols query columns from the
 *    admitted requests, always from the request with the fewest unscheduled
 *    columns, so small requests overtake large ones and the work of
 *    different requests overlaps on the device.
 *  - Graph items (params.graphs = 1): a work item of at most 256 query columns is padded to 256
 *    columns with copies of one of its non-empty columns, loaded into query
 *    buffers the slot sized once, and answered by one replay of the launch
 *    sequence (query preparation, threshold pre-pass, filtered pass,
 *    selection, re-rank with the widest list and its certificate) captured
 *    into a HIP graph the first time the search's parameters occurred; the
 *    overflow and certificate counters come back with the results, and an
 *    item that needs a second pass runs as an ordinary search instead.
 *    While a capture runs, HIP refuses synchronous work on the legacy
 *    (null) stream from every thread of the process; this library makes
 *    no such call, but a caller's own legacy-stream work on other threads
 *    can fail then.  Captures happen on the first item of each parameter set.
 *  - Results do not depend on chunking, graphs or the other requests in
 *    flight.
 * All functions may be called from any host thread. */
typedef struct sd_server sd_server;

typedef struct {
  int32_t slots;         /* concurrent workers (1..8; 0 = 4) */
  int32_t chunk_cols;    /* query columns per work item (0 = 4096) */
  int64_t budget_bytes;  /* device memory admitted requests may reserve (0 = 16 GiB) */
  int64_t max_pending;   /* requests allowed to wait for admission (0 = unlimited) */
  int32_t paused;        /* 1: workers take no work until sd_server_resume, so a batch
                            submitted meanwhile is scheduled as a whole */
  int32_t graphs;        /* 1: work items of at most 256 columns (cosine or binary,
                            precision auto or bf16) replay a HIP graph captured per search
                            parameters (see below); 0 (default): every item runs as an
                            ordinary search, which measured faster on MI355X at every
                            reference size tried */
} sd_server_params;

typedef struct {
  int64_t nrow, ncol;     /* query matrix X (CSC, host memory; copied by submit) */
  const int64_t* colptr;
  const int32_t* rowind;
  const double* values;
  sd_knn_params knn;      /* as for sd_knn(ctx, X, Y, ...); include_self is ignored */
  int32_t snn;            /* 1: also build the shared-neighbour graph of the query columns */
  int32_t borrow;         /* 1: colptr/rowind/values stay valid and unchanged until the
                             request finishes, so submit does not copy them */
  double prune;           /* SNN: keep edges with weight > prune */
} sd_request;

typedef struct {
  int32_t done;      /* 1 once the request has finished (successfully or not) */
  int32_t status;    /* SD_OK or the failure code (message: sd_server_error) */
  int32_t k_eff;     /* columns of the result */
  int32_t warn;      /* k was clamped (as sd_knn) */
  double queued_ms;  /* submit -> admission */
  double run_ms;     /* admission -> finish */
  double total_ms;   /* submit -> finish */
  int64_t finish_order; /* 1 for the first request the server finished, 2 for the next, ...
                           (0 while unfinished) */
} sd_request_info;

typedef struct {
  int64_t submitted, admitted, completed, refused;
  int64_t pending, active;  /* waiting for admission / admitted and unfinished */
  int64_t reserved_bytes;   /* current reservations of admitted requests */
  int64_t query_cols;       /* query columns of completed requests */
  int64_t packed_items;     /* work items that carried columns of several requests */
  int64_t waiting_callers;  /* callers blocked in sd_server_wait right now */
  int64_t graph_items;      /* work items answered by a graph replay */
  int64_t graph_captures;   /* graphs captured (the first item of its parameters, or after
                               a buffer the graph reads moved) */
  int64_t graph_fallbacks;  /* replays whose overflow or certificate counters sent the item
                               to an ordinary search */
  int64_t lane_admitted;    /* requests admitted into the small-request lane */
} sd_server_stats;

/* Y must stay alive until sd_server_destroy; ctx gives the device. */
int sd_server_create(sd_context* ctx, const sd_matrix* Y, const sd_server_params* params,
                     sd_server** out);
/* Queue a request; *ticket identifies it.  Validation errors of the kNN call
 * itself are reported through sd_server_wait, like any other failure. */
int sd_server_submit(sd_server* s, const sd_request* req, int64_t* ticket);
/* Wait up to timeout_s seconds (< 0: forever) for a request; fills *info. */
int sd_server_wait(sd_server* s, int64_t ticket, double timeout_s, sd_request_info* info);
/* Message of a failed request (empty if none). */
const char* sd_server_error(sd_server* s, int64_t ticket);
/* Copy a finished request's result: idx/dist hold ncol * k_eff entries in the
 * sd_knn layout; *snn (may be NULL) receives the SNN graph, whose ownership
 * passes to the caller (sd_graph_free), or NULL when none was asked for. */
int sd_server_result(sd_server* s, int64_t ticket, int32_t* idx, double* dist, sd_graph** snn);
/* Forget a finished request. */
int sd_server_release(sd_server* s, int64_t ticket);
/* Let the workers of a server created with params.paused = 1 take work. */
int sd_server_resume(sd_server* s);
int sd_server_stats_get(sd_server* s, sd_server_stats* out);
/* Stop the workers (in-flight work is cancelled) and free the server.
 * Unfinished requests end with SD_ECANCELLED; callers blocked in
 * sd_server_wait on other threads return first.  No other call on the server
 * may start once destroy has begun. */
int sd_server_destroy(sd_server* s);

/* Number of usable host cores minus one, at least 1 (R/sparseDist.R:221-224
 * over detectCoresCpp, src/detectCores.cpp:29-34).  The GPU paths ignore
 * ncores; it is kept for the R signatures. */
int sd_detect_cores(void);

#ifdef __cplusplus
}
#endif
#endif /* SD_H */
