// pair_math.h -- FP64 pair kernels with the reference's operation order.
// Shared by the exact slab, the candidate re-rank and the certificate
// fallback, so every path computes a pair value with the same code.
#pragma once

#include "sd_internal.h"

namespace sd {

// ---------------------------------------------------------------------------
// Pair kernels.  Each follows the reference's loop exactly (compiled with
// -ffp-contract=off, so no multiply-add is fused).  The query column comes
// from `q*` (LDS or global), the reference column from global memory.
// ---------------------------------------------------------------------------

// Jaccard (reference jacc_common/jacc_value, src/sparseDist.cpp:535-556).
__device__ inline double jacc_value(const int32_t* __restrict__ ra, int64_t na,
                                    const int32_t* __restrict__ rb, int64_t nb, int dist) {
  int64_t a = 0, b = 0;
  int64_t common = 0;
  while (a < na && b < nb) {
    const int32_t x = ra[a], y = rb[b];
    common += (x == y);
    a += (x <= y);
    b += (y <= x);
  }
  const double c = (double)common;
  const double denom = (double)na + (double)nb - c;
  const double sim = (denom > 0.0) ? (c / denom) : 1.0;
  return dist ? (1.0 - sim) : sim;
}

// Cosine of unit columns (reference unit_dot/unit_cosine,
// src/sparseDist.cpp:287-331): products of unit values over the common rows
// in ascending row order, empty conventions, NaN -> 0, clamp.
__device__ inline double cos_value(const int32_t* __restrict__ ra, const double* __restrict__ ua,
                                   int64_t na, double sa, const int32_t* __restrict__ rb,
                                   const double* __restrict__ ub, int64_t nb, double sb) {
  const bool ea = (sa == 0.0), eb = (sb == 0.0);
  if (ea || eb) return (ea && eb) ? 1.0 : 0.0;
  int64_t a = 0, b = 0;
  double s = 0.0;
  while (a < na && b < nb) {
    const int32_t x = ra[a], y = rb[b];
    if (x == y) {
      s += ua[a] * ub[b];
      ++a;
      ++b;
    } else if (x < y) {
      ++a;
    } else {
      ++b;
    }
  }
  if (!(s == s)) return 0.0;
  if (s > 1.0) return 1.0;
  if (s < -1.0) return -1.0;
  return s;
}

// Euclidean (reference fastEuclidean2, src/sparseDist.cpp:854-872): the sum
// of squared differences over the union of stored rows in ascending row
// order; the distance is its square root.
__device__ inline double euc_num(const int32_t* __restrict__ ra, const double* __restrict__ xa,
                                 int64_t na, const int32_t* __restrict__ rb,
                                 const double* __restrict__ xb, int64_t nb) {
  int64_t a = 0, b = 0;
  double num = 0.0;
  while (a < na && b < nb) {
    const int32_t x = ra[a], y = rb[b];
    if (x == y) {
      const double diff = xa[a] - xb[b];
      num += diff * diff;
      ++a;
      ++b;
    } else if (x < y) {
      num += xa[a] * xa[a];
      ++a;
    } else {
      num += xb[b] * xb[b];
      ++b;
    }
  }
  for (; a < na; ++a) num += xa[a] * xa[a];
  for (; b < nb; ++b) num += xb[b] * xb[b];
  return num;
}
__device__ inline double euc_value(const int32_t* __restrict__ ra, const double* __restrict__ xa,
                                   int64_t na, const int32_t* __restrict__ rb,
                                   const double* __restrict__ xb, int64_t nb) {
  return sqrt(euc_num(ra, xa, na, rb, xb, nb));
}

// Manhattan (reference fastManhattan2, src/sparseDist.cpp:750-773).
__device__ inline double man_value(const int32_t* __restrict__ ra, const double* __restrict__ xa,
                                   int64_t na, const int32_t* __restrict__ rb,
                                   const double* __restrict__ xb, int64_t nb) {
  int64_t a = 0, b = 0;
  double num = 0.0;
  while (a < na && b < nb) {
    const int32_t x = ra[a], y = rb[b];
    if (x == y) {
      num += fabs(xa[a] - xb[b]);
      ++a;
      ++b;
    } else if (x < y) {
      num += fabs(xa[a]);
      ++a;
    } else {
      num += fabs(xb[b]);
      ++b;
    }
  }
  for (; a < na; ++a) num += fabs(xa[a]);
  for (; b < nb; ++b) num += fabs(xb[b]);
  return num;
}

// Jensen-Shannon distance (reference fastJS2 and xlogx,
// src/sparseDist.cpp:206-210, 1074-1115).  Per row
// s = xlogx(p) + xlogx(q) - xlogx(p + q) + (p + q) log 2 in that order, with
// xlogx(x) = x log x and its NaN (0 log 0) taken as 0; the rows are summed
// the way Armadillo's accu sums a dense vector (even rows into one
// accumulator, odd rows into the other, then the two added), and the
// distance is sqrt(0.5 sum) when positive, else 0.  A row stored in neither
// column gives s = +0, which leaves either accumulator unchanged, so the walk
// visits only the union of stored rows.  log is the device library's FP64
// log, which may differ from the host C library's in the last bit.
__device__ inline double js_xlogx(double x) {
  const double r = x * log(x);
  return (r != r) ? 0.0 : r;
}
__device__ inline double js_value(const int32_t* __restrict__ ra, const double* __restrict__ xa,
                                  int64_t na, const int32_t* __restrict__ rb,
                                  const double* __restrict__ xb, int64_t nb) {
  const double log2 = 0x1.62e42fefa39efp-1;  // log(2), correctly rounded
  int64_t a = 0, b = 0;
  double acc[2] = {0.0, 0.0};
  while (a < na || b < nb) {
    const int32_t x = a < na ? ra[a] : INT32_MAX, y = b < nb ? rb[b] : INT32_MAX;
    const int32_t r = x < y ? x : y;
    const double p = (x == r) ? xa[a++] : 0.0;
    const double q = (y == r) ? xb[b++] : 0.0;
    const double t = js_xlogx(p) + js_xlogx(q) - js_xlogx(p + q) + (p + q) * log2;
    acc[r & 1] += t;
  }
  const double v = 0.5 * (acc[0] + acc[1]);
  return (v > 0.0) ? sqrt(v) : 0.0;
}

// Centred cross-product over all nrow rows (this project's pearson and
// covariance definition; stored-both and stored-one-side terms in ascending
// row order, then the (nrow - |union|) rows stored in neither column).
__device__ inline double centred_dot(const int32_t* __restrict__ ra, const double* __restrict__ xa,
                                     int64_t na, double ma, const int32_t* __restrict__ rb,
                                     const double* __restrict__ xb, int64_t nb, double mb,
                                     int64_t nrow) {
  int64_t a = 0, b = 0, nunion = 0;
  double s = 0.0;
  while (a < na && b < nb) {
    const int32_t x = ra[a], y = rb[b];
    if (x == y) {
      s += (xa[a] - ma) * (xb[b] - mb);
      ++a;
      ++b;
    } else if (x < y) {
      s += (xa[a] - ma) * (-mb);
      ++a;
    } else {
      s += (-ma) * (xb[b] - mb);
      ++b;
    }
    ++nunion;
  }
  for (; a < na; ++a, ++nunion) s += (xa[a] - ma) * (-mb);
  for (; b < nb; ++b, ++nunion) s += (-ma) * (xb[b] - mb);
  s += (double)(nrow - nunion) * (ma * mb);
  return s;
}

// Pearson correlation of reference column i and query column j (similarity
// form): centred cross product over the square roots of the centred sums of
// squares; non-finite -> 0, clamped to [-1, 1] (src/sparseDist.cpp:426-429).
struct ExactArgs;
__device__ inline double pear_value(const ExactArgs& A, int64_t i, const int32_t* qr,
                                    const double* qv, int64_t qn, int64_t j);

// Value of one (reference column i, query column j) pair in the reference's
// arithmetic, before any kNN masking.  `qr/qv` may point into LDS.
__device__ inline double exact_pair(const ExactArgs& A, int64_t i, const int32_t* qr, const double* qv,
                             int64_t qn, int64_t j) {
  const int64_t a0 = A.ref.p[i], na = A.ref.p[i + 1] - a0;
  const int32_t* ra = A.ref.i + a0;
  switch (A.method) {
    case SD_BINARY:
      return jacc_value(ra, na, qr, qn, A.dist);
    case SD_COSINE: {
      const double c = cos_value(ra, A.ref_unit + a0, na, A.ref_scale[i], qr, qv, qn,
                                 A.qry_scale[j]);
      const double v = A.dist ? (1.0 - c) : c;
      return (v == 0.0) ? 0.0 : v;
    }
    case SD_EUCLIDEAN:
      return euc_value(ra, A.ref.x + a0, na, qr, qv, qn);
    case SD_MANHATTAN:
      return man_value(ra, A.ref.x + a0, na, qr, qv, qn);
    case SD_JS:
      return js_value(ra, A.ref.x + a0, na, qr, qv, qn);
    case SD_PEARSON: {
      const double c = pear_value(A, i, qr, qv, qn, j);
      const double v = A.dist ? (1.0 - c) : c;
      return (v == 0.0) ? 0.0 : v;
    }
    default: {  // SD_COVARIANCE
      const double rm1 = (A.ref.nrow > 1) ? (double)(A.ref.nrow - 1) : 1.0;
      const double v = centred_dot(ra, A.ref.x + a0, na, A.ref_mean[i], qr, qv, qn,
                                   A.qry_mean[j], A.ref.nrow) /
                       rm1;
      return (v == 0.0) ? 0.0 : v;
    }
  }
}

__device__ inline double pear_value(const ExactArgs& A, int64_t i, const int32_t* qr,
                                    const double* qv, int64_t qn, int64_t j) {
  const int64_t a0 = A.ref.p[i], na = A.ref.p[i + 1] - a0;
  const double s = centred_dot(A.ref.i + a0, A.ref.x + a0, na, A.ref_mean[i], qr, qv, qn,
                               A.qry_mean[j], A.ref.nrow);
  double c = s / (sqrt(A.ref_css[i]) * sqrt(A.qry_css[j]));
  if (!isfinite(c)) c = 0.0;
  if (c > 1.0) c = 1.0;
  if (c < -1.0) c = -1.0;
  return c;
}

}  // namespace sd
