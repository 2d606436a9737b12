/*
 * sd_oracle.h -- CPU FP64 restatement of sparseDist's blocked kNN / SNN path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP library
 * (libsd_hip.so) and the reported host-CPU baseline.  Only the test suite, the
 * smoke check and the cpu_baseline leg of bench.py may load it; the product
 * path never links, loads or calls it.
 *
 * Every function restates the behaviour of a reference function (cited per
 * function in sd_oracle.c) with the same operation order and no FMA
 * contraction (built with -ffp-contract=off), so its FP64 results are the
 * reference's results bit for bit wherever the reference's arithmetic is its
 * own (cosine, euclidean, manhattan, Jaccard, top-k, SNN).  Pearson and
 * covariance call Armadillo in the reference; here they follow a definition of
 * this project's own (documented in sd_oracle.c), so their bit-level parity
 * with the reference is unpinned (tolerance parity only).
 *
 * Conventions (identical to the HIP library's C ABI):
 *   - CSC input: colptr int64[ncol+1], rowind int32[nnz] (ascending within a
 *     column), values double[nnz].
 *   - kNN output: idx int32 and dist double, nq x k column-major, 1-based
 *     indices; empty slots hold R's NA_integer_ (INT32_MIN) and NA_real_.
 *   - SNN output: an n x n CSC (dgCMatrix layout) with 0-based row indices.
 */
#ifndef SD_ORACLE_H
#define SD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OR_OK = 0,
  OR_EINVAL = 1,     /* invalid argument (message in or_last_error) */
  OR_ENONFINITE = 2, /* cosine given NA/NaN/Inf */
  OR_ENOMEM = 3,
  OR_ENOCAND = 4,    /* "No candidate neighbours available." */
  OR_ECAP = 5        /* caller buffer too small; required size reported */
};

enum {
  OR_BINARY = 0,
  OR_COSINE = 1,
  OR_EUCLIDEAN = 2,
  OR_MANHATTAN = 3,
  OR_PEARSON = 4,
  OR_COVARIANCE = 5,
  OR_JS = 6         /* Jensen-Shannon distance (finite, non-negative input) */
};

typedef struct {
  int64_t nrow, ncol;
  const int64_t* p; /* ncol + 1 */
  const int32_t* i; /* nnz */
  const double* x;  /* nnz */
} or_csc;

const char* or_last_error(void);

/* One dense slab D (m.ncol x m2.ncol, column-major): D[i + j*m.ncol] relates
 * column i of m to column j of m2. */
int or_slab(int method, const or_csc* m, const or_csc* m2, int dist, int nthreads, double* D);

/* Column-wise top-k of a dense block (reference topKBlock). */
int or_topk_block(const double* D, int64_t nrow, int64_t ncol, int k, int decreasing,
                  const int32_t* self_row, int nthreads, int32_t* idx, double* val);

/* Running row-wise top-k (reference topKRowsAccum); best_* are nrow x k and
 * are updated in place. */
int or_topk_rows_accum(const double* D, int64_t nrow, int64_t ncol, int k, int decreasing,
                       int32_t* best_idx, double* best_val, int64_t col_offset,
                       const int32_t* self_col, int nthreads, int tile_rows);

/* Blocked kNN driver (reference sparseKNN after argument validation).
 * Y == NULL means self-search.  idx/dst must hold X.ncol * k entries; the
 * result occupies the first X.ncol * (*k_eff) entries.  *warn is set to 1 when
 * k was clamped. */
int or_knn(const or_csc* X, const or_csc* Y, int k, int method, int dist, int include_self,
           int drop_degenerate, int64_t block_size, int nthreads, int32_t* idx, double* dst,
           int* k_eff, int* warn);

/* Seconds the last or_knn call on the calling thread spent in its slab
 * kernel and in its top-k reduction. */
void or_last_phases(double* slab_s, double* topk_s);

/* Both directions from one slab per block (reference sparseKNNCross).
 * warn bit 0: forward clamped, bit 1: reverse clamped.  Error OR_ENOCAND
 * reports the direction in or_last_error. */
int or_knn_cross(const or_csc* X, const or_csc* Y, int k, int method, int dist,
                 int drop_degenerate, int64_t block_size, int nthreads,
                 int32_t* fidx, double* fdst, int* k_fwd,
                 int32_t* ridx, double* rdst, int* k_rev, int* warn);

/* SNN weights over the kNN edge set (reference snnJaccard): w is n x k. */
int or_snn_jaccard(const int32_t* idx, int n, int k, int include_self, int nthreads, double* w);

/* SNN graph as an n x n CSC (reference sparseSNN).  edges: 0 = "knn",
 * 1 = "shared".  colptr has n+1 entries; rowind/val have room for cap
 * entries.  *nnz receives the entry count; OR_ECAP when cap is too small. */
int or_snn_graph(const int32_t* idx, int n, int k, int include_self, double prune,
                 int symmetrise, int edges, int nthreads, int64_t* colptr, int32_t* rowind,
                 double* val, int64_t cap, int64_t* nnz);

#ifdef __cplusplus
}
#endif
#endif
