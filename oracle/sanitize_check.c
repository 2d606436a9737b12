/*
 * sanitize_check.c -- drives every entry point of the CPU oracle on small
 * seeded inputs (empty columns, ties, NA padding, every method, both SNN edge
 * sets, buffers too small) so that `make sanitize` can run the oracle under
 * AddressSanitizer and UndefinedBehaviorSanitizer.  TEST INFRASTRUCTURE ONLY.
 * Exits non-zero when a call fails unexpectedly; the sanitizers abort on
 * their own findings.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sd_oracle.h"

static uint64_t rng_state = 88172645463325252ULL;
static uint64_t next_u64(void) {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return rng_state;
}
static double next_unit(void) { return (double)(next_u64() >> 11) / 9007199254740992.0; }

typedef struct {
  or_csc m;
  int64_t* p;
  int32_t* i;
  double* x;
} owned_csc;

/* nrow x ncol with about density * nrow stored rows per column; column
 * `empty` stores nothing; values |N|-like in (0, 1], negated when signed_. */
static owned_csc make(int64_t nrow, int64_t ncol, double density, int64_t empty, int signed_) {
  owned_csc o;
  o.p = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ncol + 1));
  o.i = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nrow * ncol + 1));
  o.x = (double*)malloc(sizeof(double) * (size_t)(nrow * ncol + 1));
  int64_t e = 0;
  o.p[0] = 0;
  for (int64_t j = 0; j < ncol; ++j) {
    for (int64_t r = 0; r < nrow; ++r) {
      if (j == empty || next_unit() >= density) continue;
      o.i[e] = (int32_t)r;
      double v = 0.05 + next_unit();
      if (signed_ && (next_u64() & 1)) v = -v;
      o.x[e++] = v;
    }
    o.p[j + 1] = e;
  }
  o.m.nrow = nrow;
  o.m.ncol = ncol;
  o.m.p = o.p;
  o.m.i = o.i;
  o.m.x = o.x;
  return o;
}

static void release(owned_csc* o) {
  free(o->p);
  free(o->i);
  free(o->x);
}

static int failures = 0;
static void expect(int rc, int want, const char* what) {
  if (rc != want) {
    fprintf(stderr, "%s: status %d (wanted %d): %s\n", what, rc, want, or_last_error());
    ++failures;
  }
}

int main(void) {
  const int64_t nrow = 40, nx = 37, ny = 23;
  owned_csc X = make(nrow, nx, 0.3, 4, 0), Y = make(nrow, ny, 0.3, 7, 0);
  owned_csc S = make(nrow, nx, 0.3, 2, 1);
  const int k = 6;
  double* D = (double*)malloc(sizeof(double) * (size_t)(nx * ny));
  int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ny * 64));
  double* dst = (double*)malloc(sizeof(double) * (size_t)(ny * 64));
  int32_t* ridx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nx * 64));
  double* rdst = (double*)malloc(sizeof(double) * (size_t)(nx * 64));
  int keff = 0, warn = 0, krev = 0;
  for (int method = OR_BINARY; method <= OR_JS; ++method) {
    const owned_csc* A = (method == OR_PEARSON || method == OR_COVARIANCE) ? &S : &X;
    for (int dist = 0; dist <= 1; ++dist) {
      const int dist_only = method == OR_EUCLIDEAN || method == OR_MANHATTAN || method == OR_JS;
      expect(or_slab(method, &A->m, &Y.m, dist, 2, D), OR_OK, "or_slab");
      for (int self = 0; self <= 1; ++self)
        expect(or_knn(&A->m, NULL, k, method, dist, self, 1, 5, 2, ridx, rdst, &keff, &warn),
               (dist_only && !dist) ? OR_EINVAL : OR_OK, "or_knn self");
      expect(or_knn(&Y.m, &A->m, k, method, dist, 0, 0, 3, 2, idx, dst, &keff, &warn),
             (dist_only && !dist) ? OR_EINVAL : OR_OK, "or_knn");
      expect(or_knn_cross(&A->m, &Y.m, 64, method, dist, 1, 4, 2, ridx, rdst, &keff, idx, dst,
                          &krev, &warn),
             (dist_only && !dist) ? OR_EINVAL : OR_OK, "or_knn_cross (k clamped)");
    }
  }
  /* top-k helpers with NaN, Inf and self rows */
  for (int64_t t = 0; t < nx * ny; ++t) D[t] = (t % 11 == 0) ? NAN : (double)(t % 5);
  D[3] = INFINITY;
  int32_t* self_row = (int32_t*)calloc((size_t)ny, sizeof(int32_t));
  self_row[1] = 2;
  for (int dec = 0; dec <= 1; ++dec)
    expect(or_topk_block(D, nx, ny, nx + 3, dec, self_row, 2, idx, dst), OR_OK, "or_topk_block");
  int32_t* bi = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nx * k));
  double* bv = (double*)malloc(sizeof(double) * (size_t)(nx * k));
  for (int64_t t = 0; t < nx * k; ++t) {
    bi[t] = INT32_MIN;
    bv[t] = NAN;
  }
  int32_t* self_col = (int32_t*)calloc((size_t)nx, sizeof(int32_t));
  self_col[0] = 1;
  expect(or_topk_rows_accum(D, nx, ny, k, 0, bi, bv, 0, self_col, 2, 8), OR_OK,
         "or_topk_rows_accum");
  /* SNN on a kNN result with NA padding, both edge sets, a buffer too small */
  const int n = (int)nx;
  expect(or_knn(&X.m, NULL, k, OR_COSINE, 1, 0, 1, 8, 2, ridx, rdst, &keff, &warn), OR_OK,
         "knn for snn");
  double* w = (double*)malloc(sizeof(double) * (size_t)(n * keff));
  for (int self = 0; self <= 1; ++self)
    expect(or_snn_jaccard(ridx, n, keff, self, 2, w), OR_OK, "or_snn_jaccard");
  int64_t* cp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
  const int64_t cap = (int64_t)n * n;
  int32_t* ri = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
  double* rv = (double*)malloc(sizeof(double) * (size_t)cap);
  int64_t nnz = 0;
  for (int edges = 0; edges <= 1; ++edges) {
    expect(or_snn_graph(ridx, n, keff, 1, 1.0 / 15, 1, edges, 2, cp, ri, rv, cap, &nnz), OR_OK,
           "or_snn_graph");
    expect(or_snn_graph(ridx, n, keff, 1, 0.0, 1, edges, 2, cp, ri, rv, 1, &nnz), OR_ECAP,
           "or_snn_graph (small buffer)");
  }
  release(&X);
  release(&Y);
  release(&S);
  free(D);
  free(idx);
  free(dst);
  free(ridx);
  free(rdst);
  free(self_row);
  free(bi);
  free(bv);
  free(self_col);
  free(w);
  free(cp);
  free(ri);
  free(rv);
  if (failures) return 1;
  printf("sanitize_check: every oracle entry point ran clean\n");
  return 0;
}
