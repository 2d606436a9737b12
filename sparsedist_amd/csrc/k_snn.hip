// k_snn.hip -- shared-nearest-neighbour graph on the device.
//
// Neighbour sets are built once per node (sorted, de-duplicated, include_self
// deciding membership; reference build_sets, src/snn.cpp:119-141).  Weights
// are integer intersection counts turned into one FP64 division each
// (reference jaccard_from_counts, src/snn.cpp:179-181).  The graph is
// assembled as (column << 32 | row) keys sorted on the device with rocPRIM,
// so the result comes out directly in dgCMatrix (CSC) order.
#include <cstring>
#include <mutex>

#include <rocprim/rocprim.hpp>

#include "sd_internal.h"

namespace sd {
namespace {

__global__ void k_validate(const int32_t* __restrict__ idx, int64_t total, int nids,
                           int* __restrict__ bad) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t v = idx[e];
    if (v != kNaInt && (v < 1 || v > nids)) atomicOr(bad, 1);
  }
}

// One thread per node: gather, insertion-sort, de-duplicate.  self_space:
// neighbour ids name the nodes themselves (a node's own id is dropped, and
// added back when include_self); otherwise they name another column space.
// Sets are rows of ks >= k + 1 entries, ks a multiple of 4 (16-byte rows).
__global__ void k_build_sets(const int32_t* __restrict__ idx, int n, int k, int ks,
                             int include_self, int self_space, int32_t* __restrict__ sets,
                             int32_t* __restrict__ len) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t* S = sets + (int64_t)i * ks;
  int m = 0;
  if (include_self && self_space) S[m++] = i;
  for (int t = 0; t < k; ++t) {
    const int32_t v = idx[(int64_t)t * n + i];
    if (v == kNaInt) continue;
    const int32_t nb = v - 1;
    if (self_space && nb == i) continue;
    int p = m++;
    while (p > 0 && S[p - 1] > nb) {
      S[p] = S[p - 1];
      --p;
    }
    S[p] = nb;
  }
  int u = 0;
  for (int a = 0; a < m; ++a)
    if (u == 0 || S[u - 1] != S[a]) S[u++] = S[a];
  len[i] = u;
}

__device__ inline int inter_size(const int32_t* a, int na, const int32_t* b, int nb) {
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

// w(i, t) over the kNN edge set (reference snnJaccard, src/snn.cpp:198-250).
// KS > 0: both sets (rows of KS entries) are read with 16-byte loads into
// registers and intersected by comparing every pair of entries (sets are
// sorted and duplicate-free, so each entry matches at most once): a row
// read per set instead of one dependent load per merge step.  KS == 0: the
// merge walk over rows of ks entries.
template <int KS>
__device__ inline double knn_weight(const int32_t* __restrict__ idx, int n, int ks,
                                    const int32_t* __restrict__ sets,
                                    const int32_t* __restrict__ len, int64_t e) {
  const int i = (int)(e % n);
  const int32_t v = idx[e];
  double out = __longlong_as_double((long long)kNaRealBits);
  if (v != kNaInt && len[i] > 0) {
    const int j = v - 1;
    const int li = len[i], lj = len[j];
    if (lj > 0) {
      int u = 0;
      if constexpr (KS > 0) {
        int32_t a[KS], b[KS];
        const int4* ri = reinterpret_cast<const int4*>(sets + (int64_t)i * KS);
        const int4* rj = reinterpret_cast<const int4*>(sets + (int64_t)j * KS);
#pragma unroll
        for (int q = 0; q < KS / 4; ++q) {
          const int4 x = ri[q], y = rj[q];
          a[4 * q] = x.x; a[4 * q + 1] = x.y; a[4 * q + 2] = x.z; a[4 * q + 3] = x.w;
          b[4 * q] = y.x; b[4 * q + 1] = y.y; b[4 * q + 2] = y.z; b[4 * q + 3] = y.w;
        }
#pragma unroll
        for (int t = 0; t < KS; ++t)
#pragma unroll
          for (int s = 0; s < KS; ++s) u += (t < li && s < lj && a[t] == b[s]) ? 1 : 0;
      } else {
        u = inter_size(sets + (int64_t)i * ks, li, sets + (int64_t)j * ks, lj);
      }
      out = (double)u / (double)(li + lj - u);
    }
  }
  return out;
}

template <int KS>
__global__ void k_knn_weights(const int32_t* __restrict__ idx, int n, int k, int ks,
                              const int32_t* __restrict__ sets, const int32_t* __restrict__ len,
                              double* __restrict__ w) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)n * k;
       e += (int64_t)gridDim.x * blockDim.x)
    w[e] = knn_weight<KS>(idx, n, ks, sets, len, e);
}

// Kept kNN entries (R/sparseSNN.R:226: !is.na(w) & !is.na(idx) & w > prune)
// as CSC keys: g(i, j) -> (j << 32 | i); symmetrise adds t(g).
__global__ void k_knn_emit(const int32_t* __restrict__ idx, const double* __restrict__ w, int n,
                           int k, double prune, int symmetrise, uint64_t* __restrict__ keys,
                           double* __restrict__ vals, unsigned long long* __restrict__ count) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)n * k;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t v = idx[e];
    const double x = w[e];
    if (v == kNaInt || isnan(x) || !(x > prune)) continue;
    const uint64_t i = (uint64_t)(e % n), j = (uint64_t)(v - 1);
    const unsigned long long at = atomicAdd(count, symmetrise ? 2ULL : 1ULL);
    keys[at] = (j << 32) | i;
    vals[at] = x;
    if (symmetrise) {
      keys[at + 1] = (i << 32) | j;
      vals[at + 1] = x;
    }
  }
}

__global__ void k_run_heads(const uint64_t* __restrict__ keys, int64_t m,
                            int64_t* __restrict__ flag) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m;
       e += (int64_t)gridDim.x * blockDim.x)
    flag[e] = (e == 0 || keys[e] != keys[e - 1]) ? 1 : 0;
}

__global__ void k_scatter_heads(const uint64_t* __restrict__ keys, const double* __restrict__ vals,
                                const int64_t* __restrict__ pos, int64_t m,
                                int32_t* __restrict__ rows, double* __restrict__ out_vals,
                                uint64_t* __restrict__ out_keys) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (e == 0 || keys[e] != keys[e - 1]) {
      const int64_t at = pos[e];
      rows[at] = (int32_t)(keys[e] & 0xFFFFFFFFu);
      out_vals[at] = vals[e];
      out_keys[at] = keys[e];
    }
  }
}

// colptr[c] = number of entries with column < c (lower bound of c << 32).
__global__ void k_colptr(const uint64_t* __restrict__ keys, int64_t m, int64_t n,
                         int64_t* __restrict__ colptr) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c > n) return;
  const uint64_t target = (uint64_t)c << 32;
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < target)
      lo = mid + 1;
    else
      hi = mid;
  }
  colptr[c] = lo;
}

// ---- shared edge set ----------------------------------------------------------
__global__ void k_rev_count(const int32_t* __restrict__ sets, const int32_t* __restrict__ len,
                            int n, int ks, int64_t* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < len[i]; ++a)
    atomicAdd((unsigned long long*)&cnt[sets[(int64_t)i * ks + a]], 1ULL);
}

__global__ void k_rev_fill(const int32_t* __restrict__ sets, const int32_t* __restrict__ len,
                           int n, int ks, const int64_t* __restrict__ rp,
                           int64_t* __restrict__ fill, int32_t* __restrict__ rv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < len[i]; ++a) {
    const int32_t m = sets[(int64_t)i * ks + a];
    const unsigned long long at = atomicAdd((unsigned long long*)&fill[m], 1ULL);
    rv[rp[m] + (int64_t)at] = i;
  }
}

__global__ void k_pair_count(const int64_t* __restrict__ rp, int n, int64_t* __restrict__ pc) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const int64_t r = rp[m + 1] - rp[m];
  pc[m] = r * (r - 1) / 2;
}

// One workgroup per m: every pair (x < y) of positions in R(m); thread x
// writes its pairs at x*(2r-x-1)/2 + (y-x-1) past the node's offset.
__global__ void k_pair_emit(const int64_t* __restrict__ rp, const int32_t* __restrict__ rv, int n,
                            const int64_t* __restrict__ poff, uint64_t* __restrict__ keys,
                            int64_t m0) {
  const int64_t m = m0 + blockIdx.x;
  const int64_t b0 = rp[m], r = rp[m + 1] - b0;
  const int64_t base = poff[m];
  for (int64_t x = threadIdx.x; x < r; x += blockDim.x) {
    const int64_t off = base + x * (2 * r - x - 1) / 2;
    const uint32_t ax = (uint32_t)rv[b0 + x];
    for (int64_t y = x + 1; y < r; ++y) {
      uint32_t a = ax, b = (uint32_t)rv[b0 + y];
      if (a > b) {
        const uint32_t t = a;
        a = b;
        b = t;
      }
      keys[off + (y - x - 1)] = ((uint64_t)a << 32) | b;
    }
  }
}

// Runs of equal (a, b) keys: run length = |N(a) n N(b)|.  Kept pairs are
// emitted in both orientations as CSC keys (the reference returns the upper
// triangle and R mirrors it, R/sparseSNN.R:206-218).
__global__ void k_shared_emit(const uint64_t* __restrict__ pk, int64_t m,
                              const int32_t* __restrict__ len, double prune,
                              uint64_t* __restrict__ keys, double* __restrict__ vals,
                              unsigned long long* __restrict__ count) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (e > 0 && pk[e] == pk[e - 1]) continue;
    int64_t end = e + 1;
    while (end < m && pk[end] == pk[e]) ++end;
    const int u = (int)(end - e);
    const uint64_t a = pk[e] >> 32, b = pk[e] & 0xFFFFFFFFu;
    const double w = (double)u / (double)(len[a] + len[b] - u);
    if (!(w > prune)) continue;
    const unsigned long long at = atomicAdd(count, 2ULL);
    keys[at] = (b << 32) | a;  // row a, column b
    vals[at] = w;
    keys[at + 1] = (a << 32) | b;
    vals[at + 1] = w;
  }
}

int sort_pairs(sd_context* ctx, DevBuf& kin, DevBuf& vin, DevBuf& kout, DevBuf& vout, int64_t m,
               int end_bit) {
  size_t bytes = 0;
  SD_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin.as<uint64_t>(), kout.as<uint64_t>(),
                                   vin.as<double>(), vout.as<double>(), (size_t)m, 0, end_bit,
                                   ctx->stream));
  DevBuf& tmp = ctx->snn_scratch[10];
  SD_TRY(tmp.ensure(bytes + 16));
  SD_HIP(rocprim::radix_sort_pairs(tmp.ptr, bytes, kin.as<uint64_t>(), kout.as<uint64_t>(),
                                   vin.as<double>(), vout.as<double>(), (size_t)m, 0, end_bit,
                                   ctx->stream));
  return SD_OK;
}

int bits_for(int64_t n) {
  int b = 1;
  while (b < 32 && ((int64_t)1 << b) < n) ++b;
  return b;
}

// Sorted (key, val) -> graph CSC with duplicates collapsed (use.last.ij:
// duplicates carry the same weight).
int assemble_csc(sd_context* ctx, DevBuf& keys, DevBuf& vals, int64_t m, int64_t n, sd_graph* g) {
  g->n = n;
  SD_TRY(g->colptr.alloc(sizeof(int64_t) * (size_t)(n + 1)));
  if (m == 0) {
    g->nnz = 0;
    SD_HIP(hipMemsetAsync(g->colptr.ptr, 0, sizeof(int64_t) * (size_t)(n + 1), ctx->stream));
    SD_TRY(g->rowind.alloc(16));
    SD_TRY(g->val.alloc(16));
    SD_HIP(hipStreamSynchronize(ctx->stream));
    return SD_OK;
  }
  DevBuf& flag = ctx->snn_scratch[11];
  DevBuf& pos = ctx->snn_scratch[12];
  SD_TRY(flag.ensure(sizeof(int64_t) * (size_t)m));
  SD_TRY(pos.ensure(sizeof(int64_t) * (size_t)(m + 1)));
  const unsigned gb = stride_grid(m);
  hipLaunchKernelGGL(k_run_heads, dim3(gb), dim3(256), 0, ctx->stream, keys.as<uint64_t>(), m,
                     flag.as<int64_t>());
  size_t bytes = 0;
  SD_HIP(rocprim::exclusive_scan(nullptr, bytes, flag.as<int64_t>(), pos.as<int64_t>(), (int64_t)0,
                                 (size_t)m, rocprim::plus<int64_t>(), ctx->stream));
  DevBuf& tmp = ctx->snn_scratch[13];
  SD_TRY(tmp.ensure(bytes + 16));
  SD_HIP(rocprim::exclusive_scan(tmp.ptr, bytes, flag.as<int64_t>(), pos.as<int64_t>(), (int64_t)0,
                                 (size_t)m, rocprim::plus<int64_t>(), ctx->stream));
  int64_t last_pos = 0, last_flag = 0;
  SD_HIP(hipMemcpyAsync(&last_pos, pos.as<int64_t>() + (m - 1), sizeof(int64_t),
                        hipMemcpyDeviceToHost, ctx->stream));
  SD_HIP(hipMemcpyAsync(&last_flag, flag.as<int64_t>() + (m - 1), sizeof(int64_t),
                        hipMemcpyDeviceToHost, ctx->stream));
  SD_HIP(hipStreamSynchronize(ctx->stream));
  const int64_t nu = last_pos + last_flag;
  g->nnz = nu;
  SD_TRY(g->rowind.alloc(sizeof(int32_t) * (size_t)nu + 16));
  SD_TRY(g->val.alloc(sizeof(double) * (size_t)nu + 16));
  DevBuf& ukeys = ctx->snn_scratch[14];
  SD_TRY(ukeys.ensure(sizeof(uint64_t) * (size_t)nu + 16));
  hipLaunchKernelGGL(k_scatter_heads, dim3(gb), dim3(256), 0, ctx->stream, keys.as<uint64_t>(),
                     vals.as<double>(), pos.as<int64_t>(), m, g->rowind.as<int32_t>(),
                     g->val.as<double>(), ukeys.as<uint64_t>());
  hipLaunchKernelGGL(k_colptr, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, ctx->stream,
                     ukeys.as<uint64_t>(), nu, n, g->colptr.as<int64_t>());
  SD_HIP(hipGetLastError());
  SD_HIP(hipStreamSynchronize(ctx->stream));
  return SD_OK;
}

// The weight kernel for rows of ks set entries (register form for ks <= 32,
// merge walk beyond).
int launch_knn_weights(sd_context* ctx, const int32_t* d_idx, int n, int k, int ks,
                       const int32_t* sets, const int32_t* len, double* dw) {
  const unsigned ge = stride_grid((int64_t)n * k);
  switch (ks) {
#define SD_WEIGHTS(KS)                                                                        \
  case KS:                                                                                    \
    hipLaunchKernelGGL(k_knn_weights<KS>, dim3(ge), dim3(256), 0, ctx->stream, d_idx, n, k, ks, \
                       sets, len, dw);                                                        \
    break;
    SD_WEIGHTS(4)
    SD_WEIGHTS(8)
    SD_WEIGHTS(12)
    SD_WEIGHTS(16)
    SD_WEIGHTS(20)
    SD_WEIGHTS(24)
    SD_WEIGHTS(28)
    SD_WEIGHTS(32)
#undef SD_WEIGHTS
    default:
      hipLaunchKernelGGL(k_knn_weights<0>, dim3(ge), dim3(256), 0, ctx->stream, d_idx, n, k, ks,
                         sets, len, dw);
  }
  SD_HIP(hipGetLastError());
  return SD_OK;
}

}  // namespace

int snn_build(sd_context* ctx, const int32_t* d_idx, int n, int k, int include_self, double prune,
              int symmetrise, int edges, sd_graph* g, double* d_w_out, int nids) {
  if (n < 1 || k < 1) return set_error(SD_EINVAL, "'idx' must have at least one row and column.");
  const bool self_space = (nids <= 0);
  if (self_space) nids = n;
  if (!self_space && edges == 0)
    return set_error(SD_EINVAL, "internal: kNN-edge weights need neighbour ids in the node space");
  const int64_t total = (int64_t)n * k;
  DevBuf& bad = ctx->snn_scratch[1];
  SD_TRY(bad.ensure(sizeof(int)));
  SD_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(int), ctx->stream));
  hipLaunchKernelGGL(k_validate, dim3(stride_grid(total)), dim3(256), 0, ctx->stream,
                     d_idx, total, nids, bad.as<int>());
  int hbad = 0;
  SD_HIP(hipMemcpyAsync(&hbad, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  SD_HIP(hipStreamSynchronize(ctx->stream));
  if (hbad)
    return set_error(SD_EINVAL, self_space
                                    ? "'idx' entries must be NA or 1-based row indices in [1, nrow(idx)]."
                                    : "neighbour ids must be NA or 1-based column indices of the reference.");

  KernelTimer timer(ctx, &ctx->stats.snn_ms, &ctx->stats.snn_launches);
  DevBuf& sets = ctx->snn_scratch[2];
  DevBuf& len = ctx->snn_scratch[3];
  const int ks = (k + 1 + 3) & ~3;  // set row stride: 16-byte rows
  SD_TRY(sets.ensure(sizeof(int32_t) * (size_t)n * (size_t)ks));
  SD_TRY(len.ensure(sizeof(int32_t) * (size_t)n));
  const unsigned gn = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_build_sets, dim3(gn), dim3(256), 0, ctx->stream, d_idx, n, k, ks,
                     include_self, (int)self_space, sets.as<int32_t>(), len.as<int32_t>());
  SD_HIP(hipGetLastError());

  const int end_bit = 32 + bits_for(n);
  if (edges == 0) {
    DevBuf& w = ctx->snn_scratch[4];
    double* dw = d_w_out;
    if (!dw) {
      SD_TRY(w.ensure(sizeof(double) * (size_t)total));
      dw = w.as<double>();
    }
    const unsigned ge = stride_grid(total);
    SD_TRY(launch_knn_weights(ctx, d_idx, n, k, ks, sets.as<int32_t>(), len.as<int32_t>(), dw));
    if (!g) {
      SD_HIP(hipStreamSynchronize(ctx->stream));
      return SD_OK;
    }
    const int64_t cap = total * (symmetrise ? 2 : 1);
    DevBuf& keys = ctx->snn_scratch[5];
    DevBuf& vals = ctx->snn_scratch[6];
    DevBuf& cnt = ctx->snn_scratch[7];
    DevBuf& keys2 = ctx->snn_scratch[8];
    DevBuf& vals2 = ctx->snn_scratch[9];
    SD_TRY(keys.ensure(sizeof(uint64_t) * (size_t)cap + 16));
    SD_TRY(vals.ensure(sizeof(double) * (size_t)cap + 16));
    SD_TRY(cnt.ensure(sizeof(unsigned long long)));
    SD_HIP(hipMemsetAsync(cnt.ptr, 0, sizeof(unsigned long long), ctx->stream));
    hipLaunchKernelGGL(k_knn_emit, dim3(ge), dim3(256), 0, ctx->stream, d_idx, dw, n, k, prune,
                       symmetrise, keys.as<uint64_t>(), vals.as<double>(),
                       cnt.as<unsigned long long>());
    unsigned long long m = 0;
    SD_HIP(hipMemcpyAsync(&m, cnt.ptr, sizeof(m), hipMemcpyDeviceToHost, ctx->stream));
    SD_HIP(hipStreamSynchronize(ctx->stream));
    if (m > 0) {
      SD_TRY(keys2.ensure(sizeof(uint64_t) * (size_t)m + 16));
      SD_TRY(vals2.ensure(sizeof(double) * (size_t)m + 16));
      SD_TRY(sort_pairs(ctx, keys, vals, keys2, vals2, (int64_t)m, end_bit));
    }
    return assemble_csc(ctx, keys2, vals2, (int64_t)m, n, g);
  }

  // ---- edges = "shared" ------------------------------------------------------
  DevBuf& rc = ctx->snn_scratch[15];
  DevBuf& rp = ctx->snn_scratch[16];
  DevBuf& fill = ctx->snn_scratch[17];
  DevBuf& rv = ctx->snn_scratch[18];
  DevBuf& pc = ctx->snn_scratch[19];
  DevBuf& poff = ctx->snn_scratch[20];
  // inverted index over the neighbour ids: R(m) = nodes whose set holds m
  const unsigned gm = (unsigned)((nids + 255) / 256);
  SD_TRY(rc.ensure(sizeof(int64_t) * (size_t)(nids + 1)));
  SD_TRY(rp.ensure(sizeof(int64_t) * (size_t)(nids + 1)));
  SD_TRY(fill.ensure(sizeof(int64_t) * (size_t)(nids + 1)));
  SD_HIP(hipMemsetAsync(rc.ptr, 0, sizeof(int64_t) * (size_t)(nids + 1), ctx->stream));
  SD_HIP(hipMemsetAsync(fill.ptr, 0, sizeof(int64_t) * (size_t)(nids + 1), ctx->stream));
  hipLaunchKernelGGL(k_rev_count, dim3(gn), dim3(256), 0, ctx->stream, sets.as<int32_t>(),
                     len.as<int32_t>(), n, ks, rc.as<int64_t>());
  size_t bytes = 0;
  SD_HIP(rocprim::exclusive_scan(nullptr, bytes, rc.as<int64_t>(), rp.as<int64_t>(), (int64_t)0,
                                 (size_t)(nids + 1), rocprim::plus<int64_t>(), ctx->stream));
  DevBuf& tmp = ctx->snn_scratch[13];
  SD_TRY(tmp.ensure(bytes + 16));
  SD_HIP(rocprim::exclusive_scan(tmp.ptr, bytes, rc.as<int64_t>(), rp.as<int64_t>(), (int64_t)0,
                                 (size_t)(nids + 1), rocprim::plus<int64_t>(), ctx->stream));
  int64_t nrev = 0;
  SD_HIP(hipMemcpyAsync(&nrev, rp.as<int64_t>() + nids, sizeof(int64_t), hipMemcpyDeviceToHost,
                        ctx->stream));
  SD_HIP(hipStreamSynchronize(ctx->stream));
  SD_TRY(rv.ensure(sizeof(int32_t) * (size_t)nrev + 16));
  hipLaunchKernelGGL(k_rev_fill, dim3(gn), dim3(256), 0, ctx->stream, sets.as<int32_t>(),
                     len.as<int32_t>(), n, ks, rp.as<int64_t>(), fill.as<int64_t>(),
                     rv.as<int32_t>());
  SD_TRY(pc.ensure(sizeof(int64_t) * (size_t)(nids + 1)));
  SD_TRY(poff.ensure(sizeof(int64_t) * (size_t)(nids + 1)));
  SD_HIP(hipMemsetAsync(pc.ptr, 0, sizeof(int64_t) * (size_t)(nids + 1), ctx->stream));
  hipLaunchKernelGGL(k_pair_count, dim3(gm), dim3(256), 0, ctx->stream, rp.as<int64_t>(), nids,
                     pc.as<int64_t>());
  bytes = 0;
  SD_HIP(rocprim::exclusive_scan(nullptr, bytes, pc.as<int64_t>(), poff.as<int64_t>(), (int64_t)0,
                                 (size_t)(nids + 1), rocprim::plus<int64_t>(), ctx->stream));
  SD_TRY(tmp.ensure(bytes + 16));
  SD_HIP(rocprim::exclusive_scan(tmp.ptr, bytes, pc.as<int64_t>(), poff.as<int64_t>(), (int64_t)0,
                                 (size_t)(nids + 1), rocprim::plus<int64_t>(), ctx->stream));
  int64_t npairs = 0;
  SD_HIP(hipMemcpyAsync(&npairs, poff.as<int64_t>() + nids, sizeof(int64_t), hipMemcpyDeviceToHost,
                        ctx->stream));
  SD_HIP(hipStreamSynchronize(ctx->stream));
  if (npairs == 0) {
    DevBuf e1, e2;
    return assemble_csc(ctx, e1, e2, 0, n, g);
  }
  DevBuf& pk = ctx->snn_scratch[21];
  DevBuf& pk2 = ctx->snn_scratch[22];
  SD_TRY(pk.ensure(sizeof(uint64_t) * (size_t)npairs + 16));
  SD_TRY(pk2.ensure(sizeof(uint64_t) * (size_t)npairs + 16));
  for (int64_t m0 = 0; m0 < nids; m0 += kMaxItemsPerLaunch) {
    const int64_t nm = std::min<int64_t>(nids - m0, kMaxItemsPerLaunch);
    hipLaunchKernelGGL(k_pair_emit, dim3((unsigned)nm), dim3(256), 0, ctx->stream,
                       rp.as<int64_t>(), rv.as<int32_t>(), nids, poff.as<int64_t>(),
                       pk.as<uint64_t>(), m0);
    SD_HIP(hipGetLastError());
  }
  bytes = 0;
  SD_HIP(rocprim::radix_sort_keys(nullptr, bytes, pk.as<uint64_t>(), pk2.as<uint64_t>(),
                                  (size_t)npairs, 0, end_bit, ctx->stream));
  SD_TRY(tmp.ensure(bytes + 16));
  SD_HIP(rocprim::radix_sort_keys(tmp.ptr, bytes, pk.as<uint64_t>(), pk2.as<uint64_t>(),
                                  (size_t)npairs, 0, end_bit, ctx->stream));
  DevBuf& keys = ctx->snn_scratch[5];
  DevBuf& vals = ctx->snn_scratch[6];
  DevBuf& cnt = ctx->snn_scratch[7];
  DevBuf& keys2 = ctx->snn_scratch[8];
  DevBuf& vals2 = ctx->snn_scratch[9];
  SD_TRY(keys.ensure(sizeof(uint64_t) * (size_t)npairs * 2 + 16));
  SD_TRY(vals.ensure(sizeof(double) * (size_t)npairs * 2 + 16));
  SD_TRY(cnt.ensure(sizeof(unsigned long long)));
  SD_HIP(hipMemsetAsync(cnt.ptr, 0, sizeof(unsigned long long), ctx->stream));
  hipLaunchKernelGGL(k_shared_emit, dim3(stride_grid(npairs)), dim3(256), 0,
                     ctx->stream, pk2.as<uint64_t>(), npairs, len.as<int32_t>(), prune,
                     keys.as<uint64_t>(), vals.as<double>(), cnt.as<unsigned long long>());
  unsigned long long m = 0;
  SD_HIP(hipMemcpyAsync(&m, cnt.ptr, sizeof(m), hipMemcpyDeviceToHost, ctx->stream));
  SD_HIP(hipStreamSynchronize(ctx->stream));
  if (m > 0) {
    SD_TRY(keys2.ensure(sizeof(uint64_t) * (size_t)m + 16));
    SD_TRY(vals2.ensure(sizeof(double) * (size_t)m + 16));
    SD_TRY(sort_pairs(ctx, keys, vals, keys2, vals2, (int64_t)m, end_bit));
  }
  return assemble_csc(ctx, keys2, vals2, (int64_t)m, n, g);
}

}  // namespace sd

// ---------------------------------------------------------------------------
// Allocation-free SNN weights (sd_snn_jaccard_noalloc).
//
// Its scratch -- the neighbour sets (n rows of ks entries), their sizes and
// the validation flag -- lives in `scratch`, one array per GPU that the
// library holds until the process ends.  In host code `scratch` only names
// the symbol: its device address on each GPU comes from hipGetSymbolAddress
// (once per device, with that device current, when a context is created) and
// only that address is ever passed to kernels.  Every call on a GPU shares
// the array, so calls that overlap on one GPU race; each call clears the
// flag it needs at zero and writes every set before reading it.
// ---------------------------------------------------------------------------
namespace sd {
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kNoAllocMaxDevices = 64;
}  // namespace sd

static __device__ __align__(256) uint8_t scratch[sd::kNoAllocScratchBytes];

namespace sd {
namespace {

std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};

__global__ void k_clear_flag(int* __restrict__ flag) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

int noalloc_scratch_setup(int device, void** out) {
  if (device < 0 || device >= kNoAllocMaxDevices)
    return set_error(SD_EINVAL, "device %d is beyond the %d this library supports", device,
                     kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(SD_EDEVICE, "scratch address on device %d: %s", device,
                       hipGetErrorString(e));
    g_noalloc_ptr[device] = p;
  }
  *out = g_noalloc_ptr[device];
  return SD_OK;
}

}  // namespace sd

extern "C" int sd_snn_jaccard_noalloc(sd_context* ctx, const int32_t* d_idx, int32_t n, int32_t k,
                                      int32_t include_self, double* d_w) {
  using namespace sd;
  clear_error();
  if (!ctx || !d_idx || !d_w) return set_error(SD_EINVAL, "NULL argument");
  if (n < 1 || k < 1) return set_error(SD_EINVAL, "'idx' must have at least one row and column.");
  const int ks = (k + 1 + 3) & ~3;  // set row stride, as snn_build
  const size_t sets_b = align256(sizeof(int32_t) * (size_t)n * (size_t)ks);
  const size_t len_b = align256(sizeof(int32_t) * (size_t)n);
  const size_t need = sets_b + len_b + 256;
  if (need > kNoAllocScratchBytes)
    return set_error(SD_EINVAL,
                     "sd_snn_jaccard_noalloc: n = %d, k = %d needs %zu bytes of scratch, more "
                     "than its %zu",
                     n, k, need, kNoAllocScratchBytes);
  if (!ctx->noalloc_scratch) return set_error(SD_EDEVICE, "context has no scratch address");
  (void)hipGetLastError();
  SD_HIP(hipSetDevice(ctx->device));
  uint8_t* base = static_cast<uint8_t*>(ctx->noalloc_scratch);
  int32_t* sets = reinterpret_cast<int32_t*>(base);
  int32_t* len = reinterpret_cast<int32_t*>(base + sets_b);
  int* flag = reinterpret_cast<int*>(base + sets_b + len_b);
  const int64_t total = (int64_t)n * k;
  // validate before any kernel reads a set through an index
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, ctx->stream, flag);
  hipLaunchKernelGGL(k_validate, dim3(stride_grid(total)), dim3(256), 0, ctx->stream, d_idx, total,
                     n, flag);
  SD_HIP(hipGetLastError());
  int hbad = 0;
  SD_HIP(hipMemcpyAsync(&hbad, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  SD_HIP(hipStreamSynchronize(ctx->stream));
  if (hbad)
    return set_error(SD_EINVAL, "'idx' entries must be NA or 1-based row indices in [1, nrow(idx)].");
  hipLaunchKernelGGL(k_build_sets, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     d_idx, n, k, ks, include_self, 1, sets, len);
  SD_HIP(hipGetLastError());
  SD_TRY(launch_knn_weights(ctx, d_idx, n, k, ks, sets, len, d_w));
  SD_HIP(hipStreamSynchronize(ctx->stream));
  return SD_OK;
}
