// k_exact.hip -- FP64 slab with the reference's operation order, and the
// exact column-wise top-k used by every path that materialises a slab
// (the exact method path, the certificate fallback and the one-to-one
// replacements of topKBlock / topKRowsAccum).
#include "sd_internal.h"
#include "pair_math.h"
#include "topk_block.h"

namespace sd {
namespace {

constexpr int kStageMax = 2048;  // query entries staged in LDS

// grid: (ceil(nref / 256), nb); thread = reference column.
__global__ __launch_bounds__(256) void k_exact_slab(ExactArgs A, const int32_t* __restrict__ qcols,
                                                    double* __restrict__ D, int64_t ld,
                                                    int mask) {
  __shared__ int32_t s_r[kStageMax];
  __shared__ double s_v[kStageMax];
  const int64_t b = blockIdx.y;
  const int64_t j = qcols ? qcols[b] : b;
  const int64_t q0 = A.qry.p[j], qn = A.qry.p[j + 1] - q0;
  const double* qvals = (A.method == SD_COSINE) ? A.qry_unit : A.qry.x;
  const int32_t* qr = A.qry.i + q0;
  const double* qv = qvals + q0;
  if (qn <= kStageMax) {
    for (int64_t t = threadIdx.x; t < qn; t += blockDim.x) {
      s_r[t] = qr[t];
      s_v[t] = qv[t];
    }
    __syncthreads();
    qr = s_r;
    qv = s_v;
  }
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= A.ref.ncol) return;
  double v;
  if (mask && A.ref_deg && A.ref_deg[i]) {
    v = __longlong_as_double((long long)kNaRealBits);
  } else if (mask && A.self_mode && i == j + A.self_off) {
    if (A.include_self) {
      // exact self value restored (R/sparseKNN.R:270-279)
      if (A.method == SD_BINARY || A.method == SD_COSINE || A.method == SD_PEARSON)
        v = A.dist ? 0.0 : 1.0;
      else if (A.method == SD_EUCLIDEAN || A.method == SD_MANHATTAN || A.method == SD_JS)
        v = 0.0;
      else
        v = exact_pair(A, i, qr, qv, qn, j);
    } else {
      v = __longlong_as_double((long long)kNaRealBits);  // excluded like self_row
    }
  } else {
    v = exact_pair(A, i, qr, qv, qn, j);
  }
  D[i + b * ld] = v;
}

// The last entry a previous pass wrote for a column: later passes admit only
// candidates ranked strictly after it ((value, index) order), so passes of
// at most kPassK entries each extend one list exactly.  An NA entry means the
// column had no more candidates.
struct PassFloor {
  bool on = false, exhausted = false;
  uint64_t key = 0;
  uint32_t id = 0;
  __device__ PassFloor(const int32_t* idx, const double* val, int decreasing) {
    if (!idx) return;
    on = true;
    if (*idx == kNaInt) {
      exhausted = true;
    } else {
      key = order_key(*val, decreasing);
      id = (uint32_t)(*idx - 1);
    }
  }
  __device__ bool admits(double v, int64_t id_, int decreasing) const {
    if (!on) return true;
    if (exhausted || !isfinite(v)) return false;
    const uint64_t kk = order_key(v, decreasing);
    return key < kk || (key == kk && id < (uint32_t)id_);
  }
};

// One workgroup per column of D: list entries [t0, t0 + k) of its top list.
template <int CAP>
__global__ __launch_bounds__(256) void k_topk_dense(const double* __restrict__ D, int64_t nrow,
                                                    int64_t ld, int k, int decreasing,
                                                    const int32_t* __restrict__ skip,
                                                    int32_t* __restrict__ out_idx,
                                                    double* __restrict__ out_val, int64_t out_ld,
                                                    const int32_t* __restrict__ outcol, int t0) {
  __shared__ BlockTopKStorage<CAP> st;
  const int64_t b = blockIdx.x;
  const double* col = D + b * ld;
  const int64_t sk = skip ? (int64_t)skip[b] : -1;
  const int64_t oc = outcol ? (int64_t)outcol[b] : b;
  const PassFloor fl(t0 > 0 ? out_idx + oc + (int64_t)(t0 - 1) * out_ld : nullptr,
                     t0 > 0 ? out_val + oc + (int64_t)(t0 - 1) * out_ld : nullptr, decreasing);
  BlockTopK<CAP> tk(st, k, decreasing);
  tk.run(nrow, [&](int64_t r, double& v, int64_t& id) {
    v = col[r];
    id = r;
    return r != sk && fl.admits(v, r, decreasing);
  });
  tk.write(out_idx + oc + (int64_t)t0 * out_ld, out_val + oc + (int64_t)t0 * out_ld, out_ld, 1);
}

// Running row-wise top-k: row r0 + blockIdx.x of D (strided by ld across
// columns) merged with its seed state (nrow x k_all, column-major, 1-based
// global indices), written as entries [t0, t0 + k) of best.  A single pass
// (k = k_all) may read its seed from best itself: every read precedes the
// writes of the same workgroup.
template <int CAP>
__global__ __launch_bounds__(256) void k_topk_rows(const double* __restrict__ D, int64_t nrow,
                                                   int64_t ncol, int k, int k_all, int decreasing,
                                                   const int32_t* seed_idx, const double* seed_val,
                                                   int32_t* best_idx, double* best_val,
                                                   int64_t col_offset,
                                                   const int32_t* __restrict__ self_col,
                                                   int64_t r0, int t0) {
  __shared__ BlockTopKStorage<CAP> st;
  const int64_t r = r0 + blockIdx.x;
  const PassFloor fl(t0 > 0 ? best_idx + (int64_t)(t0 - 1) * nrow + r : nullptr,
                     t0 > 0 ? best_val + (int64_t)(t0 - 1) * nrow + r : nullptr, decreasing);
  BlockTopK<CAP> tk(st, k, decreasing);
  // seed: slots fill left to right until the first NA (src/topKRows.cpp:187-199)
  int nseed = 0;
  while (nseed < k_all) {
    const int32_t gi = seed_idx[(int64_t)nseed * nrow + r];
    if (gi == kNaInt || gi <= 0) break;
    ++nseed;
  }
  const int64_t sc = self_col[r];
  tk.run(nseed + ncol, [&](int64_t t, double& v, int64_t& id) {
    if (t < nseed) {
      v = seed_val[t * nrow + r];
      id = (int64_t)seed_idx[t * nrow + r] - 1;
      return fl.admits(v, id, decreasing);
    }
    const int64_t b = t - nseed;
    v = D[r + b * nrow];
    id = col_offset + b;
    return (id + 1) != sc && fl.admits(v, id, decreasing);
  });
  __syncthreads();  // every seed read of this row precedes the writes below
  tk.write(best_idx + (int64_t)t0 * nrow + r, best_val + (int64_t)t0 * nrow + r, nrow, 1);
}

}  // namespace

int exact_slab(sd_context* ctx, const ExactArgs& a, const int32_t* d_qcols, int64_t nb, double* D,
               int64_t ld, int mask) {
  if (nb <= 0 || a.ref.ncol <= 0) return SD_OK;
  const int64_t chunk = 65535;  // grid.y limit
  for (int64_t b0 = 0; b0 < nb; b0 += chunk) {
    const int64_t n = (nb - b0 < chunk) ? nb - b0 : chunk;
    dim3 grid((unsigned)((a.ref.ncol + 255) / 256), (unsigned)n);
    hipLaunchKernelGGL(k_exact_slab, grid, dim3(256), 0, ctx->stream, a,
                       d_qcols ? d_qcols + b0 : nullptr, D + b0 * ld, ld, mask);
    SD_HIP(hipGetLastError());
  }
  return SD_OK;
}

// List entries one top-k pass keeps (BlockTopK holds k + 256 in LDS).
constexpr int kPassK = 1792;

static int cap_for(int k) {
  if (k + 256 <= 512) return 512;
  if (k + 256 <= 1024) return 1024;
  return 2048;
}

int topk_dense(sd_context* ctx, const double* D, int64_t nrow, int64_t ncol, int64_t ld, int k,
               int decreasing, const int32_t* d_skip, int32_t* out_idx, double* out_val,
               int64_t out_ld, const int32_t* d_outcol) {
  if (ncol <= 0 || k <= 0) return SD_OK;
  // k above one pass's capacity: passes of kPassK entries, each after the
  // last entry of the one before
  for (int t0 = 0; t0 < k; t0 += kPassK) {
    const int kc = (k - t0 < kPassK) ? k - t0 : kPassK;
    const int cap = cap_for(kc);
    for (int64_t c0 = 0; c0 < ncol; c0 += kMaxItemsPerLaunch) {
      const int64_t n = (ncol - c0 < kMaxItemsPerLaunch) ? ncol - c0 : kMaxItemsPerLaunch;
#define SD_LAUNCH_TOPK(C)                                                                        \
  hipLaunchKernelGGL(k_topk_dense<C>, dim3((unsigned)n), dim3(256), 0, ctx->stream, D + c0 * ld, \
                     nrow, ld, kc, decreasing, d_skip ? d_skip + c0 : nullptr,                   \
                     d_outcol ? out_idx : out_idx + c0, d_outcol ? out_val : out_val + c0, out_ld, \
                     d_outcol ? d_outcol + c0 : nullptr, t0)
      switch (cap) {
        case 512: SD_LAUNCH_TOPK(512); break;
        case 1024: SD_LAUNCH_TOPK(1024); break;
        default: SD_LAUNCH_TOPK(2048); break;
      }
#undef SD_LAUNCH_TOPK
      SD_HIP(hipGetLastError());
    }
  }
  return SD_OK;
}

int topk_rows(sd_context* ctx, const double* D, int64_t nrow, int64_t ncol, int k, int decreasing,
              int32_t* best_idx, double* best_val, int64_t col_offset, const int32_t* self_col) {
  if (nrow <= 0 || k <= 0) return SD_OK;
  // several passes rewrite best while later passes still need the seed:
  // they read it from a copy
  const int32_t* seed_idx = best_idx;
  const double* seed_val = best_val;
  if (k > kPassK) {
    const size_t n = (size_t)nrow * (size_t)k;
    SD_TRY(ctx->scratch[18].ensure(sizeof(int32_t) * n + 64));
    SD_TRY(ctx->scratch[19].ensure(sizeof(double) * n + 64));
    SD_HIP(hipMemcpyAsync(ctx->scratch[18].ptr, best_idx, sizeof(int32_t) * n,
                          hipMemcpyDeviceToDevice, ctx->stream));
    SD_HIP(hipMemcpyAsync(ctx->scratch[19].ptr, best_val, sizeof(double) * n,
                          hipMemcpyDeviceToDevice, ctx->stream));
    seed_idx = ctx->scratch[18].as<int32_t>();
    seed_val = ctx->scratch[19].as<double>();
  }
  for (int t0 = 0; t0 < k; t0 += kPassK) {
    const int kc = (k - t0 < kPassK) ? k - t0 : kPassK;
    const int cap = cap_for(kc);
    for (int64_t r0 = 0; r0 < nrow; r0 += kMaxItemsPerLaunch) {
      const int64_t n = (nrow - r0 < kMaxItemsPerLaunch) ? nrow - r0 : kMaxItemsPerLaunch;
#define SD_LAUNCH_ROWS(C)                                                                      \
  hipLaunchKernelGGL(k_topk_rows<C>, dim3((unsigned)n), dim3(256), 0, ctx->stream, D, nrow, ncol, \
                     kc, k, decreasing, seed_idx, seed_val, best_idx, best_val, col_offset,       \
                     self_col, r0, t0)
      switch (cap) {
        case 512: SD_LAUNCH_ROWS(512); break;
        case 1024: SD_LAUNCH_ROWS(1024); break;
        default: SD_LAUNCH_ROWS(2048); break;
      }
#undef SD_LAUNCH_ROWS
      SD_HIP(hipGetLastError());
    }
  }
  return SD_OK;
}

// A degenerate query column gets no neighbours: NA index and NA distance in
// every slot of its row (R/sparseKNN.R:298-301).
__global__ void k_fill_degenerate(int32_t* __restrict__ idx, double* __restrict__ val,
                                  const uint8_t* __restrict__ deg, int64_t nq, int64_t n) {
  const double na = __longlong_as_double((long long)kNaRealBits);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (deg[e % nq]) {
      idx[e] = kNaInt;
      val[e] = na;
    }
  }
}

int fill_degenerate(sd_context* ctx, int32_t* idx, double* val, const uint8_t* deg, int64_t nq,
                    int k) {
  const int64_t n = nq * k;
  if (n <= 0) return SD_OK;
  hipLaunchKernelGGL(k_fill_degenerate, dim3(stride_grid(n)), dim3(256), 0, ctx->stream, idx, val,
                     deg, nq, n);
  SD_HIP(hipGetLastError());
  return SD_OK;
}

}  // namespace sd
