/*
 * sd_oracle.c -- CPU FP64 restatement of sparseDist's blocked kNN / SNN path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the reported CPU baseline); the
 * product library never uses it.  Built with -O2 -ffp-contract=off so that no
 * multiply-add is fused: every FP64 operation below happens in the order and
 * association written, which is the order of the reference source cited next
 * to it (paths relative to the reference repository root).
 */
#include "sd_oracle.h"

#include <math.h>
#include <time.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NA_INT INT32_MIN

static __thread char g_err[512];

const char* or_last_error(void) { return g_err; }

static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

static double na_real(void) {
  /* R's NA_real_: a NaN whose low word is 1954. */
  const uint64_t bits = 0x7FF00000000007A2ULL;
  double d;
  memcpy(&d, &bits, sizeof d);
  return d;
}

static int nthreads_of(int n) { return n < 1 ? 1 : n; }

/* Seconds the last or_knn call on this thread spent in its slab kernel and
 * in its top-k (the rest of its time is driver work: zero dropping, masks,
 * copies), for the phase split of the CPU baseline. */
static __thread double g_slab_s, g_topk_s;
static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
void or_last_phases(double* slab_s, double* topk_s) {
  *slab_s = g_slab_s;
  *topk_s = g_topk_s;
}

/* ------------------------------------------------------------------------ */
/* Per-column helpers                                                        */
/* ------------------------------------------------------------------------ */

/* Unit-normalised values and per-column max |v| (reference UnitCsc/unit_csc,
 * src/sparseDist.cpp:235-284): scaled sum-of-squares recurrence in stored
 * order, then (v / scale) / sqrt(ssq).  Rejects non-finite input. */
static int unit_csc(const or_csc* m, double** vals_out, double** scale_out) {
  const int64_t nc = m->ncol;
  const int64_t nnz = m->p[nc];
  double* vals = (double*)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
  double* scl = (double*)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
  if (!vals || !scl) {
    free(vals);
    free(scl);
    return fail(OR_ENOMEM, "out of memory");
  }
  for (int64_t j = 0; j < nc; ++j) {
    double scale = 0.0, ssq = 1.0;
    for (int64_t e = m->p[j]; e < m->p[j + 1]; ++e) {
      const double v = m->x[e];
      if (!isfinite(v)) {
        free(vals);
        free(scl);
        return fail(OR_ENONFINITE, "Cosine requires finite input (found NA, NaN or Inf).");
      }
      if (v == 0.0) continue;
      const double a = fabs(v);
      if (scale < a) {
        const double r = scale / a;
        ssq = 1.0 + ssq * r * r;
        scale = a;
      } else {
        const double r = a / scale;
        ssq += r * r;
      }
    }
    const double sq = sqrt(ssq);
    scl[j] = scale;
    for (int64_t e = m->p[j]; e < m->p[j + 1]; ++e)
      vals[e] = (scale == 0.0) ? 0.0 : ((m->x[e]) / scale) / sq;
  }
  *vals_out = vals;
  *scale_out = scl;
  return OR_OK;
}

/* Column mean and centred sum of squares over all nrow entries (stored and
 * implicit zeros).  Pearson and covariance call Armadillo's cor()/cov() in the
 * reference (src/sparseDist.cpp:425, 503); this project defines them itself:
 *   mean = (sum of stored values in stored order) / nrow
 *   ss   = sum over stored of (x - mean)^2, then + (nrow - nnz) * (mean*mean). */
static void centre_stats(const or_csc* m, double* mean, double* ss) {
  const double r = (double)m->nrow;
  for (int64_t j = 0; j < m->ncol; ++j) {
    double s = 0.0;
    for (int64_t e = m->p[j]; e < m->p[j + 1]; ++e) s += m->x[e];
    const double mu = s / r;
    double q = 0.0;
    for (int64_t e = m->p[j]; e < m->p[j + 1]; ++e) {
      const double d = m->x[e] - mu;
      q += d * d;
    }
    q += (double)(m->nrow - (m->p[j + 1] - m->p[j])) * (mu * mu);
    mean[j] = mu;
    ss[j] = q;
  }
}

/* ------------------------------------------------------------------------ */
/* Pair kernels                                                              */
/* ------------------------------------------------------------------------ */

/* Jaccard (reference jacc_common/jacc_value, src/sparseDist.cpp:535-556). */
static double jacc_pair(const or_csc* a, int64_t i, const or_csc* b, int64_t j, int dist) {
  int64_t x = a->p[i], xe = a->p[i + 1], y = b->p[j], ye = b->p[j + 1];
  const double ni = (double)(xe - x), nj = (double)(ye - y);
  double common = 0.0;
  while (x < xe && y < ye) {
    const int32_t ra = a->i[x], rb = b->i[y];
    common += (double)(ra == rb);
    x += (ra <= rb);
    y += (rb <= ra);
  }
  const double denom = ni + nj - common;
  const double sim = (denom > 0.0) ? (common / denom) : 1.0;
  return dist ? (1.0 - sim) : sim;
}

/* Cosine of unit columns (reference unit_dot/unit_cosine,
 * src/sparseDist.cpp:287-331). */
static double cos_pair(const or_csc* a, const double* ua, const double* sa, int64_t i,
                       const or_csc* b, const double* ub, const double* sb, int64_t j) {
  const int ea = (sa[i] == 0.0), eb = (sb[j] == 0.0);
  if (ea || eb) return (ea && eb) ? 1.0 : 0.0;
  int64_t x = a->p[i], xe = a->p[i + 1], y = b->p[j], ye = b->p[j + 1];
  double s = 0.0;
  while (x < xe && y < ye) {
    if (a->i[x] == b->i[y]) {
      s += ua[x] * ub[y];
      ++x;
      ++y;
    } else if (a->i[x] < b->i[y]) {
      ++x;
    } else {
      ++y;
    }
  }
  if (!(s == s)) return 0.0;
  if (s > 1.0) return 1.0;
  if (s < -1.0) return -1.0;
  return s;
}

/* Euclidean (reference fastEuclidean2, src/sparseDist.cpp:854-872). */
static double euc_pair(const or_csc* a, int64_t i, const or_csc* b, int64_t j) {
  int64_t x = a->p[i], xe = a->p[i + 1], y = b->p[j], ye = b->p[j + 1];
  double num = 0.0;
  while (x < xe && y < ye) {
    if (a->i[x] == b->i[y]) {
      const double diff = a->x[x] - b->x[y];
      num += diff * diff;
      ++x;
      ++y;
    } else if (a->i[x] < b->i[y]) {
      num += a->x[x] * a->x[x];
      ++x;
    } else {
      num += b->x[y] * b->x[y];
      ++y;
    }
  }
  for (; x < xe; ++x) num += a->x[x] * a->x[x];
  for (; y < ye; ++y) num += b->x[y] * b->x[y];
  return sqrt(num);
}

/* Manhattan (reference fastManhattan2, src/sparseDist.cpp:750-773). */
static double man_pair(const or_csc* a, int64_t i, const or_csc* b, int64_t j) {
  int64_t x = a->p[i], xe = a->p[i + 1], y = b->p[j], ye = b->p[j + 1];
  double num = 0.0;
  while (x < xe && y < ye) {
    if (a->i[x] == b->i[y]) {
      num += fabs(a->x[x] - b->x[y]);
      ++x;
      ++y;
    } else if (a->i[x] < b->i[y]) {
      num += fabs(a->x[x]);
      ++x;
    } else {
      num += fabs(b->x[y]);
      ++y;
    }
  }
  for (; x < xe; ++x) num += fabs(a->x[x]);
  for (; y < ye; ++y) num += fabs(b->x[y]);
  return num;
}

/* Centred cross-product over all nrow rows by a merge walk: stored-both,
 * stored-one-side terms in ascending row order, then the rows stored in
 * neither column as one (nrow - |union|) * (mx*my) term. */
static double centred_dot(const or_csc* a, int64_t i, double mx, const or_csc* b, int64_t j,
                          double my) {
  int64_t x = a->p[i], xe = a->p[i + 1], y = b->p[j], ye = b->p[j + 1];
  double s = 0.0;
  int64_t nunion = 0;
  while (x < xe && y < ye) {
    if (a->i[x] == b->i[y]) {
      s += (a->x[x] - mx) * (b->x[y] - my);
      ++x;
      ++y;
    } else if (a->i[x] < b->i[y]) {
      s += (a->x[x] - mx) * (-my);
      ++x;
    } else {
      s += (-mx) * (b->x[y] - my);
      ++y;
    }
    ++nunion;
  }
  for (; x < xe; ++x, ++nunion) s += (a->x[x] - mx) * (-my);
  for (; y < ye; ++y, ++nunion) s += (-mx) * (b->x[y] - my);
  s += (double)(a->nrow - nunion) * (mx * my);
  return s;
}

/* Jensen-Shannon distance (reference fastJS2 and xlogx,
 * src/sparseDist.cpp:206-210, 1074-1115), over the columns as the dense
 * vectors the reference takes: per row s = xlogx(p) + xlogx(q) - xlogx(p + q)
 * + (p + q) log 2, evaluated left to right, with xlogx(x) = x log x and its
 * NaN (0 log 0) replaced by 0; the rows are summed as Armadillo's accu of a
 * dense vector does (arrayops::accumulate: even positions into one
 * accumulator, odd positions into another, then their sum); the distance is
 * sqrt(0.5 sum) when that is positive, else 0.  Armadillo is not vendored
 * and its version is not pinned, so this summation order is a restatement of
 * its published code, not a reference-pinned fact. */
static double xlogx(double x) {
  const double r = x * log(x);
  return (r != r) ? 0.0 : r;
}

static double js_pair(const or_csc* a, int64_t i, const or_csc* b, int64_t j) {
  const double log2 = log(2.0);
  int64_t x = a->p[i], xe = a->p[i + 1], y = b->p[j], ye = b->p[j + 1];
  double acc1 = 0.0, acc2 = 0.0;
  for (int64_t r = 0; r < a->nrow; ++r) {
    double p = 0.0, q = 0.0;
    if (x < xe && a->i[x] == r) p = a->x[x++];
    if (y < ye && b->i[y] == r) q = b->x[y++];
    const double s = xlogx(p) + xlogx(q) - xlogx(p + q) + (p + q) * log2;
    if (r & 1)
      acc2 += s;
    else
      acc1 += s;
  }
  const double v = 0.5 * (acc1 + acc2);
  return (v > 0.0) ? sqrt(v) : 0.0;
}

static int js_check(const or_csc* m) {
  const int64_t nnz = m->p[m->ncol];
  for (int64_t e = 0; e < nnz; ++e)
    if (!isfinite(m->x[e]) || m->x[e] < 0.0)
      return fail(OR_EINVAL,
                  "fastJS2 requires finite, non-negative input (columns should be normalised to "
                  "sum to 1 for a true JS divergence).");
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Slab                                                                      */
/* ------------------------------------------------------------------------ */

int or_slab(int method, const or_csc* m, const or_csc* m2, int dist, int nthreads, double* D) {
  if (m->nrow != m2->nrow)
    return fail(OR_EINVAL, "Mismatched row dimensions of m (%lld) and m2 (%lld).",
                (long long)m->nrow, (long long)m2->nrow);
  const int64_t nc = m->ncol, nc2 = m2->ncol;
  const int nt = nthreads_of(nthreads);
  double *ua = NULL, *sa = NULL, *ub = NULL, *sb = NULL;
  double *ma = NULL, *qa = NULL, *mb = NULL, *qb = NULL;
  int rc;
  if (method == OR_COSINE) {
    if ((rc = unit_csc(m, &ua, &sa)) != OR_OK) return rc;
    if ((rc = unit_csc(m2, &ub, &sb)) != OR_OK) {
      free(ua);
      free(sa);
      return rc;
    }
  } else if (method == OR_PEARSON || method == OR_COVARIANCE) {
    ma = (double*)malloc(sizeof(double) * (size_t)(nc + 1));
    qa = (double*)malloc(sizeof(double) * (size_t)(nc + 1));
    mb = (double*)malloc(sizeof(double) * (size_t)(nc2 + 1));
    qb = (double*)malloc(sizeof(double) * (size_t)(nc2 + 1));
    if (!ma || !qa || !mb || !qb) {
      free(ma);
      free(qa);
      free(mb);
      free(qb);
      return fail(OR_ENOMEM, "out of memory");
    }
    centre_stats(m, ma, qa);
    centre_stats(m2, mb, qb);
  } else if (method == OR_JS) {
    if ((rc = js_check(m)) != OR_OK || (rc = js_check(m2)) != OR_OK) return rc;
  } else if (method != OR_BINARY && method != OR_EUCLIDEAN && method != OR_MANHATTAN) {
    return fail(OR_EINVAL, "unknown method %d", method);
  }
  const double rm1 = (m->nrow > 1) ? (double)(m->nrow - 1) : 1.0;

#pragma omp parallel for num_threads(nt) schedule(dynamic)
  for (int64_t i = 0; i < nc; ++i) {
    for (int64_t j = 0; j < nc2; ++j) {
      double v;
      switch (method) {
        case OR_BINARY:
          v = jacc_pair(m, i, m2, j, dist);
          break;
        case OR_COSINE: {
          const double c = cos_pair(m, ua, sa, i, m2, ub, sb, j);
          v = dist ? (1.0 - c) : c;
          break;
        }
        case OR_EUCLIDEAN:
          v = euc_pair(m, i, m2, j);
          break;
        case OR_MANHATTAN:
          v = man_pair(m, i, m2, j);
          break;
        case OR_JS:
          v = js_pair(m, i, m2, j);
          break;
        case OR_PEARSON: {
          const double s = centred_dot(m, i, ma[i], m2, j, mb[j]);
          double c = s / (sqrt(qa[i]) * sqrt(qb[j]));
          if (!isfinite(c)) c = 0.0;
          if (c > 1.0) c = 1.0;
          if (c < -1.0) c = -1.0;
          v = dist ? (1.0 - c) : c;
          break;
        }
        default: /* OR_COVARIANCE */
          v = centred_dot(m, i, ma[i], m2, j, mb[j]) / rm1;
          break;
      }
      /* Similarity / coefficient slabs come back from the reference as sparse
       * matrices (src/sparseDist.cpp:438-440, 509) and are densified in R: a
       * stored -0.0 cannot survive that round trip. */
      if (v == 0.0) v = 0.0;
      D[i + j * nc] = v;
    }
  }
  free(ua);
  free(sa);
  free(ub);
  free(sb);
  free(ma);
  free(qa);
  free(mb);
  free(qb);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Bounded heaps with the (value, index) order, lower index winning ties     */
/* (reference Better/BetterRow, src/topK.cpp:42-49, src/topKRows.cpp:55-62). */
/* ------------------------------------------------------------------------ */

typedef struct {
  double v;
  int64_t i;
} cand;

static inline int better(const cand* a, const cand* b, int dec) {
  if (a->v != b->v) return dec ? (a->v > b->v) : (a->v < b->v);
  return a->i < b->i;
}

/* Max-heap under "worse": h[0] is the worst kept candidate. */
static void sift_down(cand* h, int n, int pos, int dec) {
  for (;;) {
    int l = 2 * pos + 1, r = l + 1, w = pos;
    if (l < n && better(&h[w], &h[l], dec)) w = l;
    if (r < n && better(&h[w], &h[r], dec)) w = r;
    if (w == pos) return;
    cand t = h[w];
    h[w] = h[pos];
    h[pos] = t;
    pos = w;
  }
}

static void sift_up(cand* h, int pos, int dec) {
  while (pos > 0) {
    int par = (pos - 1) / 2;
    if (!better(&h[par], &h[pos], dec)) return;
    cand t = h[par];
    h[par] = h[pos];
    h[pos] = t;
    pos = par;
  }
}

static void heap_offer(cand* h, int* n, int K, cand c, int dec) {
  if (*n < K) {
    h[*n] = c;
    sift_up(h, *n, dec);
    ++*n;
  } else if (better(&c, &h[0], dec)) {
    h[0] = c;
    sift_down(h, K, 0, dec);
  }
}

/* Sort best first (reference sort_heap). */
static void heap_sort_best_first(cand* h, int n, int dec) {
  for (int end = n - 1; end > 0; --end) {
    cand t = h[0];
    h[0] = h[end];
    h[end] = t;
    sift_down(h, end, 0, dec);
  }
  /* h is now worst-last reversed: after popping, h[0..n) is ascending in
   * "worse", i.e. best first. */
}

int or_topk_block(const double* D, int64_t nrow, int64_t ncol, int k, int decreasing,
                  const int32_t* self_row, int nthreads, int32_t* idx, double* val) {
  /* reference topKBlock, src/topK.cpp:67-149 */
  if (k < 1) return fail(OR_EINVAL, "k must be >= 1.");
  for (int64_t b = 0; b < ncol; ++b) {
    const int32_t sr = self_row[b];
    if (sr == NA_INT || sr < 0 || (int64_t)sr > nrow)
      return fail(OR_EINVAL, "self_row entries must be 0 or valid 1-based rows of D.");
  }
  const double NA = na_real();
  const int nt = nthreads_of(nthreads);
  int rc = OR_OK;
#pragma omp parallel num_threads(nt)
  {
    cand* h = (cand*)malloc(sizeof(cand) * (size_t)(k + 1));
    if (!h) {
#pragma omp atomic write
      rc = OR_ENOMEM;
    }
#pragma omp for schedule(dynamic)
    for (int64_t b = 0; b < ncol; ++b) {
      if (!h) continue;
      const int64_t skip = self_row[b] > 0 ? (int64_t)self_row[b] - 1 : nrow;
      int n = 0;
      const double* col = D + b * nrow;
      for (int64_t i = 0; i < nrow; ++i) {
        if (i == skip) continue;
        const double v = col[i];
        if (!isfinite(v)) continue;
        cand c = {v, i};
        heap_offer(h, &n, k, c, decreasing);
      }
      heap_sort_best_first(h, n, decreasing);
      for (int t = 0; t < k; ++t) {
        const int64_t at = (int64_t)t * ncol + b;
        if (t < n) {
          val[at] = h[t].v;
          idx[at] = (int32_t)(h[t].i + 1);
        } else {
          val[at] = NA;
          idx[at] = NA_INT;
        }
      }
    }
    free(h);
  }
  if (rc != OR_OK) return fail(rc, "out of memory");
  return OR_OK;
}

int or_topk_rows_accum(const double* D, int64_t nrow, int64_t ncol, int k, int decreasing,
                       int32_t* best_idx, double* best_val, int64_t col_offset,
                       const int32_t* self_col, int nthreads, int tile_rows) {
  /* reference topKRowsAccum, src/topKRows.cpp:104-248 */
  if (k < 1) return fail(OR_EINVAL, "k must be >= 1.");
  if (col_offset < 0) return fail(OR_EINVAL, "col_offset must be >= 0.");
  if (tile_rows < 1) tile_rows = 1;
  for (int64_t i = 0; i < nrow; ++i) {
    if (self_col[i] == NA_INT || self_col[i] < 0)
      return fail(OR_EINVAL, "self_col entries must be 0 or a positive 1-based query index.");
  }
  const double NA = na_real();
  const int nt = nthreads_of(nthreads);
  const int64_t TR = tile_rows;
  const int64_t ntiles = (nrow + TR - 1) / TR;
  int rc = OR_OK;
#pragma omp parallel num_threads(nt)
  {
    cand* hb = (cand*)malloc(sizeof(cand) * (size_t)TR * (size_t)(k + 1));
    int* hl = (int*)malloc(sizeof(int) * (size_t)TR);
    if (!hb || !hl) {
#pragma omp atomic write
      rc = OR_ENOMEM;
    }
#pragma omp for schedule(dynamic)
    for (int64_t tt = 0; tt < ntiles; ++tt) {
      if (!hb || !hl) continue;
      const int64_t r0 = tt * TR;
      const int64_t r1 = (r0 + TR < nrow) ? r0 + TR : nrow;
      const int64_t nr = r1 - r0;
      for (int64_t r = 0; r < nr; ++r) {
        cand* h = hb + r * (k + 1);
        int n = 0;
        const int64_t i = r0 + r;
        for (int t = 0; t < k; ++t) {
          const int64_t at = (int64_t)t * nrow + i;
          const int32_t gi = best_idx[at];
          if (gi == NA_INT || gi <= 0) break;
          cand c = {best_val[at], (int64_t)gi - 1};
          heap_offer(h, &n, k, c, decreasing);
        }
        hl[r] = n;
      }
      for (int64_t b = 0; b < ncol; ++b) {
        const int64_t g = col_offset + b;
        const double* col = D + b * nrow;
        for (int64_t r = 0; r < nr; ++r) {
          const int64_t i = r0 + r;
          if ((int64_t)self_col[i] == g + 1) continue;
          const double v = col[i];
          if (!isfinite(v)) continue;
          cand c = {v, g};
          heap_offer(hb + r * (k + 1), &hl[r], k, c, decreasing);
        }
      }
      for (int64_t r = 0; r < nr; ++r) {
        cand* h = hb + r * (k + 1);
        const int n = hl[r];
        heap_sort_best_first(h, n, decreasing);
        const int64_t i = r0 + r;
        for (int t = 0; t < k; ++t) {
          const int64_t at = (int64_t)t * nrow + i;
          if (t < n) {
            best_val[at] = h[t].v;
            best_idx[at] = (int32_t)(h[t].i + 1);
          } else {
            best_val[at] = NA;
            best_idx[at] = NA_INT;
          }
        }
      }
    }
    free(hb);
    free(hl);
  }
  if (rc != OR_OK) return fail(rc, "out of memory");
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Drivers                                                                   */
/* ------------------------------------------------------------------------ */

/* colSums(abs(M)) == 0 (reference R/sparseKNN.R:214-217): every stored value
 * is exactly zero; a NaN makes the sum NA, which is not "empty". */
static void col_is_empty(const or_csc* m, unsigned char* out) {
  for (int64_t j = 0; j < m->ncol; ++j) {
    int empty = 1;
    for (int64_t e = m->p[j]; e < m->p[j + 1]; ++e)
      if (!(m->x[e] == 0.0)) {
        empty = 0;
        break;
      }
    out[j] = (unsigned char)empty;
  }
}

/* The dgCMatrix -> arma::sp_mat conversion at the reference's .Call boundary
 * (src/RcppExports.cpp) stores no explicit zeros; drop them here so the
 * stored structure Jaccard counts is the same.  *owned receives allocations
 * to free (NULL when the input had none). */
static int drop_zeros(const or_csc* in, or_csc* out, void** owned) {
  *out = *in;
  *owned = NULL;
  const int64_t nnz = in->p[in->ncol];
  int64_t nz = 0;
  for (int64_t e = 0; e < nnz; ++e) nz += (in->x[e] == 0.0);
  if (nz == 0) return OR_OK;
  const int64_t keep = nnz - nz;
  size_t bytes = sizeof(int64_t) * (size_t)(in->ncol + 1) + sizeof(int32_t) * (size_t)(keep + 1) +
                 sizeof(double) * (size_t)(keep + 1) + 16;
  char* buf = (char*)malloc(bytes);
  if (!buf) return fail(OR_ENOMEM, "out of memory");
  int64_t* p = (int64_t*)buf;
  double* x = (double*)(p + in->ncol + 1);
  int32_t* ri = (int32_t*)(x + keep + 1);
  int64_t w = 0;
  p[0] = 0;
  for (int64_t j = 0; j < in->ncol; ++j) {
    for (int64_t e = in->p[j]; e < in->p[j + 1]; ++e)
      if (in->x[e] != 0.0) {
        ri[w] = in->i[e];
        x[w] = in->x[e];
        ++w;
      }
    p[j + 1] = w;
  }
  out->p = p;
  out->i = ri;
  out->x = x;
  *owned = buf;
  return OR_OK;
}

static int method_decreasing(int method, int dist, int* dec) {
  /* reference R/sparseKNN.R:168-181 */
  switch (method) {
    case OR_BINARY:
    case OR_COSINE:
    case OR_PEARSON:
      *dec = !dist;
      return OR_OK;
    case OR_COVARIANCE:
      *dec = 1;
      return OR_OK;
    case OR_EUCLIDEAN:
    case OR_MANHATTAN:
    case OR_JS:
      if (!dist)
        return fail(OR_EINVAL,
                    "'dist = FALSE' is only meaningful for methods \"binary\", \"cosine\" and "
                    "\"pearson\"; %s has no similarity form.",
                    method == OR_EUCLIDEAN ? "euclidean" : method == OR_JS ? "js" : "manhattan");
      *dec = 0;
      return OR_OK;
    default:
      return fail(OR_EINVAL, "unknown method %d", method);
  }
}

/* Columns [from, to) of m with colptr rebased to 0 (the slab kernels index
 * p[] relative to the arrays they are given). */
static int col_block(const or_csc* m, int64_t from, int64_t to, int64_t* pbuf, or_csc* out) {
  const int64_t base = m->p[from];
  for (int64_t j = from; j <= to; ++j) pbuf[j - from] = m->p[j] - base;
  out->nrow = m->nrow;
  out->ncol = to - from;
  out->p = pbuf;
  out->i = m->i + base;
  out->x = m->x + base;
  return OR_OK;
}

int or_knn(const or_csc* Xin, const or_csc* Yin, int k, int method, int dist, int include_self,
           int drop_degenerate, int64_t block_size, int nthreads, int32_t* idx, double* dst,
           int* k_eff, int* warn) {
  /* reference sparseKNN, R/sparseKNN.R:110-313 (after argument validation) */
  *warn = 0;
  g_slab_s = g_topk_s = 0.0;
  if (k < 1) return fail(OR_EINVAL, "'k' must be a positive integer no larger than .Machine$integer.max.");
  if (block_size < 1)
    return fail(OR_EINVAL, "'block_size' must be a positive integer no larger than .Machine$integer.max.");
  int dec, rc;
  if ((rc = method_decreasing(method, dist, &dec)) != OR_OK) return rc;
  const int selfmode = (Yin == NULL);
  if (Xin->ncol < 1) return fail(OR_EINVAL, "'X' must have at least one column.");
  if (!selfmode && Xin->nrow != Yin->nrow)
    return fail(OR_EINVAL, "'X' and 'Y' must have the same number of rows.");

  or_csc X, Y;
  void *ox = NULL, *oy = NULL;
  if ((rc = drop_zeros(Xin, &X, &ox)) != OR_OK) return rc;
  if (!selfmode && (rc = drop_zeros(Yin, &Y, &oy)) != OR_OK) {
    free(ox);
    return rc;
  }
  const or_csc* ref = selfmode ? &X : &Y;
  const int64_t nq = X.ncol, nref = ref->ncol;

  unsigned char* deg_ref = (unsigned char*)calloc((size_t)nref + 1, 1);
  unsigned char* deg_q = selfmode ? deg_ref : (unsigned char*)calloc((size_t)nq + 1, 1);
  /* the original (zero-carrying) columns decide emptiness, as colSums does */
  if (drop_degenerate) {
    col_is_empty(selfmode ? Xin : Yin, deg_ref);
    if (!selfmode) col_is_empty(Xin, deg_q);
  }
  int64_t avail = 0;
  for (int64_t j = 0; j < nref; ++j) avail += !deg_ref[j];
  if (selfmode && !include_self) avail -= 1;
  if (avail < 1) {
    rc = fail(OR_ENOCAND, "No candidate neighbours available.");
    goto out;
  }
  if (k > avail) {
    *warn = 1;
    k = (int)avail;
  }
  *k_eff = k;

  {
    const double NA = na_real();
    const int64_t bs = block_size < nq ? block_size : nq;
    double* D = (double*)malloc(sizeof(double) * (size_t)nref * (size_t)bs);
    int32_t* self_row = (int32_t*)malloc(sizeof(int32_t) * (size_t)bs);
    int32_t* pidx = (int32_t*)malloc(sizeof(int32_t) * (size_t)bs * (size_t)k);
    double* pval = (double*)malloc(sizeof(double) * (size_t)bs * (size_t)k);
    int64_t* pb = (int64_t*)malloc(sizeof(int64_t) * (size_t)(bs + 1));
    if (!D || !self_row || !pidx || !pval || !pb) {
      rc = fail(OR_ENOMEM, "out of memory");
    } else {
      for (int64_t from = 0; from < nq && rc == OR_OK; from += block_size) {
        const int64_t to = (from + block_size < nq) ? from + block_size : nq;
        const int64_t nb = to - from;
        or_csc blk;
        col_block(&X, from, to, pb, &blk);
        double t0 = now_s();
        rc = or_slab(method, ref, &blk, dist, nthreads, D);
        g_slab_s += now_s() - t0;
        if (rc != OR_OK) break;
        for (int64_t b = 0; b < nb; ++b)
          self_row[b] = (selfmode && !include_self) ? (int32_t)(from + b + 1) : 0;
        if (selfmode && include_self) {
          /* exact self value restored (R/sparseKNN.R:270-279) */
          for (int64_t b = 0; b < nb; ++b) {
            double* at = &D[(from + b) + b * nref];
            if (method == OR_BINARY || method == OR_COSINE || method == OR_PEARSON)
              *at = dist ? 0.0 : 1.0;
            else if (method == OR_EUCLIDEAN || method == OR_MANHATTAN || method == OR_JS)
              *at = 0.0;
          }
        }
        for (int64_t i = 0; i < nref; ++i)
          if (deg_ref[i])
            for (int64_t b = 0; b < nb; ++b) D[i + b * nref] = NA;
        t0 = now_s();
        rc = or_topk_block(D, nref, nb, k, dec, self_row, nthreads, pidx, pval);
        g_topk_s += now_s() - t0;
        if (rc != OR_OK) break;
        for (int t = 0; t < k; ++t)
          for (int64_t b = 0; b < nb; ++b) {
            idx[(int64_t)t * nq + from + b] = pidx[(int64_t)t * nb + b];
            dst[(int64_t)t * nq + from + b] = pval[(int64_t)t * nb + b];
          }
      }
      if (rc == OR_OK)
        for (int64_t q = 0; q < nq; ++q)
          if (deg_q[q])
            for (int t = 0; t < k; ++t) {
              idx[(int64_t)t * nq + q] = NA_INT;
              dst[(int64_t)t * nq + q] = NA;
            }
    }
    free(D);
    free(self_row);
    free(pidx);
    free(pval);
    free(pb);
  }
out:
  if (!selfmode) free(deg_q);
  free(deg_ref);
  free(ox);
  free(oy);
  return rc;
}

int or_knn_cross(const or_csc* Xin, const or_csc* Yin, int k, int method, int dist,
                 int drop_degenerate, int64_t block_size, int nthreads, int32_t* fidx,
                 double* fdst, int* k_fwd, int32_t* ridx, double* rdst, int* k_rev, int* warn) {
  /* reference sparseKNNCross, R/sparseKNNCross.R:44-199 (after validation) */
  *warn = 0;
  if (k < 1) return fail(OR_EINVAL, "'k' must be a positive integer no larger than .Machine$integer.max.");
  if (block_size < 1)
    return fail(OR_EINVAL, "'block_size' must be a positive integer no larger than .Machine$integer.max.");
  int dec, rc;
  if ((rc = method_decreasing(method, dist, &dec)) != OR_OK) return rc;
  if (Xin->nrow != Yin->nrow) return fail(OR_EINVAL, "'X' and 'Y' must have the same number of rows.");
  if (Xin->ncol < 1) return fail(OR_EINVAL, "'X' must have at least one column.");
  if (Yin->ncol < 1) return fail(OR_EINVAL, "'Y' must have at least one column.");

  or_csc X, Y;
  void *ox = NULL, *oy = NULL;
  if ((rc = drop_zeros(Xin, &X, &ox)) != OR_OK) return rc;
  if ((rc = drop_zeros(Yin, &Y, &oy)) != OR_OK) {
    free(ox);
    return rc;
  }
  const int64_t nq = X.ncol, nr = Y.ncol;
  unsigned char* deg_q = (unsigned char*)calloc((size_t)nq + 1, 1);
  unsigned char* deg_r = (unsigned char*)calloc((size_t)nr + 1, 1);
  if (drop_degenerate) {
    col_is_empty(Xin, deg_q);
    col_is_empty(Yin, deg_r);
  }
  int64_t av_f = 0, av_r = 0;
  for (int64_t j = 0; j < nr; ++j) av_f += !deg_r[j];
  for (int64_t j = 0; j < nq; ++j) av_r += !deg_q[j];
  int kf = k, kr = k;
  if (av_f < 1) {
    rc = fail(OR_ENOCAND, "No candidate neighbours available for the forward direction.");
    goto out;
  }
  if (kf > av_f) {
    *warn |= 1;
    kf = (int)av_f;
  }
  if (av_r < 1) {
    rc = fail(OR_ENOCAND, "No candidate neighbours available for the reverse direction.");
    goto out;
  }
  if (kr > av_r) {
    *warn |= 2;
    kr = (int)av_r;
  }
  *k_fwd = kf;
  *k_rev = kr;
  {
    const double NA = na_real();
    for (int64_t e = 0; e < nq * (int64_t)kf; ++e) {
      fidx[e] = NA_INT;
      fdst[e] = NA;
    }
    for (int64_t e = 0; e < nr * (int64_t)kr; ++e) {
      ridx[e] = NA_INT;
      rdst[e] = NA;
    }
    const int64_t bs = block_size < nq ? block_size : nq;
    double* D = (double*)malloc(sizeof(double) * (size_t)nr * (size_t)bs);
    int32_t* zeros_b = (int32_t*)calloc((size_t)bs, sizeof(int32_t));
    int32_t* zeros_r = (int32_t*)calloc((size_t)nr, sizeof(int32_t));
    int32_t* pidx = (int32_t*)malloc(sizeof(int32_t) * (size_t)bs * (size_t)kf);
    double* pval = (double*)malloc(sizeof(double) * (size_t)bs * (size_t)kf);
    int64_t* pb = (int64_t*)malloc(sizeof(int64_t) * (size_t)(bs + 1));
    if (!D || !zeros_b || !zeros_r || !pidx || !pval || !pb) {
      rc = fail(OR_ENOMEM, "out of memory");
    } else {
      for (int64_t from = 0; from < nq && rc == OR_OK; from += block_size) {
        const int64_t to = (from + block_size < nq) ? from + block_size : nq;
        const int64_t nb = to - from;
        or_csc blk;
        col_block(&X, from, to, pb, &blk);
        rc = or_slab(method, &Y, &blk, dist, nthreads, D);
        if (rc != OR_OK) break;
        for (int64_t i = 0; i < nr; ++i)
          if (deg_r[i])
            for (int64_t b = 0; b < nb; ++b) D[i + b * nr] = NA;
        for (int64_t b = 0; b < nb; ++b)
          if (deg_q[from + b])
            for (int64_t i = 0; i < nr; ++i) D[i + b * nr] = NA;
        rc = or_topk_block(D, nr, nb, kf, dec, zeros_b, nthreads, pidx, pval);
        if (rc != OR_OK) break;
        for (int t = 0; t < kf; ++t)
          for (int64_t b = 0; b < nb; ++b) {
            fidx[(int64_t)t * nq + from + b] = pidx[(int64_t)t * nb + b];
            fdst[(int64_t)t * nq + from + b] = pval[(int64_t)t * nb + b];
          }
        rc = or_topk_rows_accum(D, nr, nb, kr, dec, ridx, rdst, from, zeros_r, nthreads, 64);
      }
      if (rc == OR_OK) {
        for (int64_t q = 0; q < nq; ++q)
          if (deg_q[q])
            for (int t = 0; t < kf; ++t) {
              fidx[(int64_t)t * nq + q] = NA_INT;
              fdst[(int64_t)t * nq + q] = NA;
            }
        for (int64_t q = 0; q < nr; ++q)
          if (deg_r[q])
            for (int t = 0; t < kr; ++t) {
              ridx[(int64_t)t * nr + q] = NA_INT;
              rdst[(int64_t)t * nr + q] = NA;
            }
      }
    }
    free(D);
    free(zeros_b);
    free(zeros_r);
    free(pidx);
    free(pval);
    free(pb);
  }
out:
  free(deg_q);
  free(deg_r);
  free(ox);
  free(oy);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* SNN                                                                       */
/* ------------------------------------------------------------------------ */

static int cmp_i32(const void* a, const void* b) {
  const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* Validate idx (reference copy_indices, src/snn.cpp:94-108) and build every
 * neighbour set once: ascending, de-duplicated, 0-based; include_self decides
 * membership (reference build_sets, src/snn.cpp:119-141).  Sets live in one
 * (n x (k+1)) buffer with lengths in len[]. */
static int build_sets(const int32_t* idx, int n, int k, int include_self, int32_t** sets_out,
                      int* len) {
  if (n < 1 || k < 1) return fail(OR_EINVAL, "'idx' must have at least one row and column.");
  const int64_t N = n;
  for (int t = 0; t < k; ++t)
    for (int i = 0; i < n; ++i) {
      const int32_t v = idx[i + N * t];
      if (v != NA_INT && (v < 1 || v > n))
        return fail(OR_EINVAL, "'idx' entries must be NA or 1-based row indices in [1, nrow(idx)].");
    }
  int32_t* sets = (int32_t*)malloc(sizeof(int32_t) * (size_t)N * (size_t)(k + 1));
  if (!sets) return fail(OR_ENOMEM, "out of memory");
  for (int i = 0; i < n; ++i) {
    int32_t* S = sets + (int64_t)i * (k + 1);
    int m = 0;
    if (include_self) S[m++] = i;
    for (int t = 0; t < k; ++t) {
      const int32_t v = idx[i + N * t];
      if (v == NA_INT) continue;
      if (v - 1 != i) S[m++] = v - 1;
    }
    qsort(S, (size_t)m, sizeof(int32_t), cmp_i32);
    int u = 0;
    for (int a = 0; a < m; ++a)
      if (u == 0 || S[u - 1] != S[a]) S[u++] = S[a];
    len[i] = u;
  }
  *sets_out = sets;
  return OR_OK;
}

static int inter_size(const int32_t* a, int na, const int32_t* b, int nb) {
  /* reference inter_size, src/snn.cpp:73-82 */
  int c = 0, x = 0, y = 0;
  while (x < na && y < nb) {
    if (a[x] < b[y])
      ++x;
    else if (b[y] < a[x])
      ++y;
    else {
      ++c;
      ++x;
      ++y;
    }
  }
  return c;
}

static double jaccard_from_counts(int u, int ni, int nj) {
  /* reference jaccard_from_counts, src/snn.cpp:179-181 */
  return (double)u / (double)(ni + nj - u);
}

int or_snn_jaccard(const int32_t* idx, int n, int k, int include_self, int nthreads, double* w) {
  /* reference snnJaccard, src/snn.cpp:198-250 */
  int32_t* sets;
  int* len = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!len) return fail(OR_ENOMEM, "out of memory");
  int rc = build_sets(idx, n, k, include_self, &sets, len);
  if (rc != OR_OK) {
    free(len);
    return rc;
  }
  const double NA = na_real();
  const int64_t N = n;
  const int nt = nthreads_of(nthreads);
#pragma omp parallel for num_threads(nt) schedule(dynamic)
  for (int i = 0; i < n; ++i) {
    for (int t = 0; t < k; ++t) w[i + N * t] = NA;
    if (len[i] == 0) continue;
    const int32_t* Si = sets + (int64_t)i * (k + 1);
    for (int t = 0; t < k; ++t) {
      const int32_t v = idx[i + N * t];
      if (v == NA_INT) continue;
      const int j = v - 1;
      if (len[j] == 0) continue;
      const int u = inter_size(Si, len[i], sets + (int64_t)j * (k + 1), len[j]);
      w[i + N * t] = jaccard_from_counts(u, len[i], len[j]);
    }
  }
  free(sets);
  free(len);
  return OR_OK;
}

typedef struct {
  int32_t col, row;
  double v;
  int32_t src; /* 0 = g, 1 = t(g) */
} trip;

static int cmp_trip(const void* a, const void* b) {
  const trip* x = (const trip*)a;
  const trip* y = (const trip*)b;
  if (x->col != y->col) return (x->col > y->col) - (x->col < y->col);
  if (x->row != y->row) return (x->row > y->row) - (x->row < y->row);
  return (x->src > y->src) - (x->src < y->src);
}

int or_snn_graph(const int32_t* idx, int n, int k, int include_self, double prune, int symmetrise,
                 int edges, int nthreads, int64_t* colptr, int32_t* rowind, double* val,
                 int64_t cap, int64_t* nnz_out) {
  /* reference sparseSNN, R/sparseSNN.R:120-251 */
  if (!(prune >= 0.0 && prune <= 1.0))
    return fail(OR_EINVAL, "'prune' must be a finite number between 0 and 1.");
  const int64_t N = n;
  trip* T = NULL;
  int64_t nt_ = 0;
  int rc;
  if (edges == 0) {
    double* w = (double*)malloc(sizeof(double) * (size_t)N * (size_t)(k > 0 ? k : 1));
    if (!w) return fail(OR_ENOMEM, "out of memory");
    if ((rc = or_snn_jaccard(idx, n, k, include_self, nthreads, w)) != OR_OK) {
      free(w);
      return rc;
    }
    T = (trip*)malloc(sizeof(trip) * (size_t)(2 * N * k + 1));
    if (!T) {
      free(w);
      return fail(OR_ENOMEM, "out of memory");
    }
    /* keep <- !is.na(w) & !is.na(idx) & (w > prune)  (R/sparseSNN.R:226);
     * duplicates of (i, j) keep the last (use.last.ij, same weight). */
    for (int64_t i = 0; i < N; ++i)
      for (int t = 0; t < k; ++t) {
        const double x = w[i + N * t];
        const int32_t v = idx[i + N * t];
        if (isnan(x) || v == NA_INT || !(x > prune)) continue;
        trip a = {v - 1, (int32_t)i, x, 0};
        T[nt_++] = a;
        if (symmetrise) {
          trip b = {(int32_t)i, v - 1, x, 1};
          T[nt_++] = b;
        }
      }
    free(w);
    qsort(T, (size_t)nt_, sizeof(trip), cmp_trip);
    /* Combine: entries of g (src 0, duplicates collapse) and of t(g) (src 1)
     * at one position give g + t(g) - t(g) * (g > 0) (R/sparseSNN.R:241-248). */
    int64_t w_ = 0;
    for (int64_t a = 0; a < nt_;) {
      int64_t b = a;
      int has_g = 0, has_t = 0;
      double gv = 0.0, tv = 0.0;
      while (b < nt_ && T[b].col == T[a].col && T[b].row == T[a].row) {
        if (T[b].src == 0) {
          has_g = 1;
          gv = T[b].v;
        } else {
          has_t = 1;
          tv = T[b].v;
        }
        ++b;
      }
      double v;
      if (!symmetrise)
        v = gv;
      else {
        const double s = gv + tv;
        v = s - tv * (has_g ? 1.0 : 0.0);
      }
      (void)has_t;
      T[w_] = T[a];
      T[w_].v = v;
      ++w_;
      a = b;
    }
    nt_ = w_;
  } else {
    /* edges = "shared": reference snnJaccardShared, src/snn.cpp:299-447.
     * Inverted index R(m) as CSR; per row i a dense accumulator over j > i. */
    int32_t* sets;
    int* len = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (!len) return fail(OR_ENOMEM, "out of memory");
    if ((rc = build_sets(idx, n, k, include_self, &sets, len)) != OR_OK) {
      free(len);
      return rc;
    }
    int64_t* rp = (int64_t*)calloc((size_t)N + 1, sizeof(int64_t));
    int64_t tot = 0;
    for (int64_t i = 0; i < N; ++i) tot += len[i];
    int32_t* rv = (int32_t*)malloc(sizeof(int32_t) * (size_t)(tot + 1));
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(N + 1));
    int* acc = (int*)calloc((size_t)N + 1, sizeof(int));
    int32_t* touched = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N + 1));
    int64_t tcap = 1024;
    T = (trip*)malloc(sizeof(trip) * (size_t)tcap);
    if (!rp || !rv || !fill || !acc || !touched || !T) {
      free(rp);
      free(rv);
      free(fill);
      free(acc);
      free(touched);
      free(T);
      free(sets);
      free(len);
      return fail(OR_ENOMEM, "out of memory");
    }
    for (int64_t i = 0; i < N; ++i)
      for (int a = 0; a < len[i]; ++a) ++rp[sets[i * (k + 1) + a] + 1];
    for (int64_t m = 0; m < N; ++m) rp[m + 1] += rp[m];
    for (int64_t m = 0; m < N; ++m) fill[m] = rp[m];
    for (int64_t i = 0; i < N; ++i)
      for (int a = 0; a < len[i]; ++a) rv[fill[sets[i * (k + 1) + a]]++] = (int32_t)i;
    for (int64_t i = 0; i < N; ++i) {
      if (len[i] == 0) continue;
      int nt2 = 0;
      const int32_t* Si = sets + i * (k + 1);
      for (int a = 0; a < len[i]; ++a) {
        const int64_t m = Si[a];
        for (int64_t q = rp[m]; q < rp[m + 1]; ++q) {
          const int32_t j = rv[q];
          if (j <= i) continue;
          if (acc[j]++ == 0) touched[nt2++] = j;
        }
      }
      for (int a = 0; a < nt2; ++a) {
        const int32_t j = touched[a];
        const double w = jaccard_from_counts(acc[j], len[i], len[j]);
        acc[j] = 0;
        if (w > prune) {
          if (nt_ + 2 > tcap) {
            tcap *= 2;
            trip* T2 = (trip*)realloc(T, sizeof(trip) * (size_t)tcap);
            if (!T2) {
              free(rp);
              free(rv);
              free(fill);
              free(acc);
              free(touched);
              free(T);
              free(sets);
              free(len);
              return fail(OR_ENOMEM, "out of memory");
            }
            T = T2;
          }
          trip x = {j, (int32_t)i, w, 0};
          trip y = {(int32_t)i, j, w, 0};
          T[nt_++] = x;
          T[nt_++] = y;
        }
      }
    }
    free(rp);
    free(rv);
    free(fill);
    free(acc);
    free(touched);
    free(sets);
    free(len);
    qsort(T, (size_t)nt_, sizeof(trip), cmp_trip);
  }
  *nnz_out = nt_;
  if (nt_ > cap) {
    free(T);
    return fail(OR_ECAP, "output capacity %lld < %lld entries", (long long)cap, (long long)nt_);
  }
  for (int64_t j = 0; j <= N; ++j) colptr[j] = 0;
  for (int64_t e = 0; e < nt_; ++e) {
    ++colptr[T[e].col + 1];
    rowind[e] = T[e].row;
    val[e] = T[e].v;
  }
  for (int64_t j = 0; j < N; ++j) colptr[j + 1] += colptr[j];
  free(T);
  return OR_OK;
}
