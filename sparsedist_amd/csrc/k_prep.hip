// k_prep.hip -- per-column statistics and densified layouts of a resident
// CSC matrix.  Everything here runs once per matrix and is cached with it.
#include <hip/hip_bf16.h>

#include "sd_internal.h"

namespace sd {
namespace {

// Cosine unit normalisation, one thread per column (reference UnitCsc /
// unit_csc, src/sparseDist.cpp:235-284).  The scaled sum-of-squares
// recurrence is order dependent in FP64, so it runs sequentially over the
// column's stored values, exactly as the reference does; the division by
// scale and then by sqrt(ssq) is elementwise and runs in a second kernel.
__global__ void k_unit_scale(const int64_t* __restrict__ p, const double* __restrict__ x,
                             int64_t ncol, double* __restrict__ scale_out,
                             double* __restrict__ sq_out) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  double scale = 0.0, ssq = 1.0;
  for (int64_t e = p[j], ee = p[j + 1]; e < ee; ++e) {
    const double v = x[e];
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
  scale_out[j] = scale;
  sq_out[j] = sqrt(ssq);
}

// One wave per column: u = (v / scale) / sqrt(ssq), 0 for an all-zero column.
__global__ void k_unit_vals(const int64_t* __restrict__ p, const double* __restrict__ x,
                            int64_t ncol, const double* __restrict__ scale,
                            const double* __restrict__ sq, double* __restrict__ u) {
  const int lane = threadIdx.x & 63;
  const int64_t j = blockIdx.x * (int64_t)(blockDim.x / 64) + (threadIdx.x >> 6);
  if (j >= ncol) return;
  const double s = scale[j], q = sq[j];
  for (int64_t e = p[j] + lane, ee = p[j + 1]; e < ee; e += 64)
    u[e] = (s == 0.0) ? 0.0 : (x[e] / s) / q;
}

// Column mean and centred sum of squares over all nrow entries, sequential
// in stored order (this project's pearson / covariance definition; the
// reference calls Armadillo's cor()/cov(), src/sparseDist.cpp:425, 503).
__global__ void k_centre(const int64_t* __restrict__ p, const double* __restrict__ x, int64_t ncol,
                         int64_t nrow, double* __restrict__ mean, double* __restrict__ css) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  const double r = (double)nrow;
  double s = 0.0;
  for (int64_t e = p[j], ee = p[j + 1]; e < ee; ++e) s += x[e];
  const double mu = s / r;
  double q = 0.0;
  for (int64_t e = p[j], ee = p[j + 1]; e < ee; ++e) {
    const double d = x[e] - mu;
    q += d * d;
  }
  q += (double)(nrow - (p[j + 1] - p[j])) * (mu * mu);
  mean[j] = mu;
  css[j] = q;
}

// Sum of squares per column (euclidean candidate filter only; not on the
// exact path, so the order of summation is free).
__global__ void k_sqnorm(const int64_t* __restrict__ p, const double* __restrict__ x, int64_t ncol,
                         double* __restrict__ out) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  double s = 0.0;
  for (int64_t e = p[j], ee = p[j + 1]; e < ee; ++e) s += x[e] * x[e];
  out[j] = s;
}

__device__ inline uint16_t f32_to_bf16_rne(float f) {
  uint32_t b = __float_as_uint(f);
  const uint32_t lsb = (b >> 16) & 1u;
  b += 0x7FFFu + lsb;  // round to nearest even (inputs are finite)
  return (uint16_t)(b >> 16);
}

// Densify column chunks into the chunk-major layout of dense_offset (128-byte
// K chunks, each a block of ncol columns).  One workgroup per (column, K
// range): the range is zeroed in LDS, the column's stored entries that fall
// into it are scattered, and it is written out with 16-byte stores, eight
// lanes per 128-byte chunk.  ELEM bytes per element; ELEM == 0 is the FP4
// pattern layout: two rows per byte, row 2i in the low nibble, stored
// entries as e2m1 1.0 (0x2), set with LDS atomics since neighbouring rows
// share a byte.
template <int KIND, int ELEM>
__global__ __launch_bounds__(256) void k_densify(const int64_t* __restrict__ p,
                                                 const int32_t* __restrict__ ri,
                                                 const double* __restrict__ vals, int64_t ncol,
                                                 int64_t kpad, int64_t chunk,
                                                 uint8_t* __restrict__ out) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int64_t nchunks = (kpad + chunk - 1) / chunk;
  const int64_t j = blockIdx.x / nchunks;
  const int64_t c = blockIdx.x % nchunks;
  if (j >= ncol) return;
  const int64_t k0 = c * chunk;
  const int64_t len = (kpad - k0 < chunk) ? (kpad - k0) : chunk;  // multiple of 64
  const int64_t bytes = ELEM ? len * ELEM : len / 2;
  uint4* l4 = reinterpret_cast<uint4*>(lds);
  for (int64_t t = threadIdx.x; t < bytes / 16; t += blockDim.x) l4[t] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  for (int64_t e = p[j] + threadIdx.x, ee = p[j + 1]; e < ee; e += blockDim.x) {
    const int64_t r = ri[e];
    if (r < k0 || r >= k0 + len) continue;
    const int64_t o = r - k0;
    if (KIND == kDenseF4Pattern) {
      atomicOr(reinterpret_cast<uint32_t*>(lds) + (o >> 3), 0x2u << ((o & 7) * 4));
    } else if (KIND == kDenseBf16Unit || KIND == kDenseBf16Raw) {
      reinterpret_cast<uint16_t*>(lds)[o] = f32_to_bf16_rne((float)vals[e]);
    } else if (KIND == kDenseF32Pattern) {
      reinterpret_cast<float*>(lds)[o] = 1.0f;
    } else {
      reinterpret_cast<float*>(lds)[o] = (float)vals[e];
    }
  }
  __syncthreads();
  // chunk-major store: each 128-byte chunk of the column to its own block
  const int64_t kb0 = ELEM ? k0 * ELEM : k0 / 2;  // a multiple of 128
  for (int64_t t = threadIdx.x; t < bytes / 16; t += blockDim.x)
    *reinterpret_cast<uint4*>(out + dense_offset(j, kb0 + 16 * t, ncol)) = l4[t];
}

// Split layout: one workgroup per (column, K chunk) as above; LDS holds the
// chunk's hi values followed by its lo values, and the chunk is written to
// the row's three segments [hi | hi | lo] (each kpad elements).  v - hi is
// exact in FP64 (hi is v rounded to 8 significant bits, so the difference
// has at most 45 significant bits), then rounded like hi.
__global__ __launch_bounds__(256) void k_densify_split(const int64_t* __restrict__ p,
                                                       const int32_t* __restrict__ ri,
                                                       const double* __restrict__ vals,
                                                       int64_t ncol, int64_t kpad, int64_t chunk,
                                                       uint8_t* __restrict__ out) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int64_t nchunks = (kpad + chunk - 1) / chunk;
  const int64_t j = blockIdx.x / nchunks;
  const int64_t c = blockIdx.x % nchunks;
  if (j >= ncol) return;
  const int64_t k0 = c * chunk;
  const int64_t len = (kpad - k0 < chunk) ? (kpad - k0) : chunk;  // multiple of 256
  uint16_t* hi = reinterpret_cast<uint16_t*>(lds);
  uint16_t* lo = hi + len;
  uint4* l4 = reinterpret_cast<uint4*>(lds);
  for (int64_t t = threadIdx.x; t < len * 4 / 16; t += blockDim.x) l4[t] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  for (int64_t e = p[j] + threadIdx.x, ee = p[j + 1]; e < ee; e += blockDim.x) {
    const int64_t r = ri[e];
    if (r < k0 || r >= k0 + len) continue;
    const double v = vals[e];
    const uint16_t h = f32_to_bf16_rne((float)v);
    const double hv = (double)__uint_as_float((uint32_t)h << 16);
    hi[r - k0] = h;
    lo[r - k0] = f32_to_bf16_rne((float)(v - hv));
  }
  __syncthreads();
  const int64_t seg = kpad * 2;  // bytes
  const int64_t kb0 = k0 * 2;    // K byte of the chunk in the first segment (a multiple of 128)
  const int64_t n16 = len * 2 / 16;
  const uint4* h4 = reinterpret_cast<const uint4*>(hi);
  const uint4* o4 = reinterpret_cast<const uint4*>(lo);
  for (int64_t t = threadIdx.x; t < n16; t += blockDim.x) {
    const uint4 hv = h4[t];
    const int64_t kb = kb0 + 16 * t;
    *reinterpret_cast<uint4*>(out + dense_offset(j, kb, ncol)) = hv;
    *reinterpret_cast<uint4*>(out + dense_offset(j, kb + seg, ncol)) = hv;
    *reinterpret_cast<uint4*>(out + dense_offset(j, kb + 2 * seg, ncol)) = o4[t];
  }
}

// FP32 pearson aux: sqrt(nrow) mean and 1 / sqrt(centred sum of squares)
// (0 when the sum is 0: a constant column, whose correlations are all 0).
__global__ void k_centre_f32(const double* __restrict__ mean, const double* __restrict__ css,
                             int64_t ncol, int64_t nrow, float* __restrict__ cm,
                             float* __restrict__ ci) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  cm[j] = (float)(sqrt((double)nrow) * mean[j]);
  ci[j] = css[j] > 0.0 ? (float)(1.0 / sqrt(css[j])) : 0.f;
}

// FP32 copies of the FP64 squared norms.
__global__ void k_to_f32(const double* __restrict__ in, int64_t n, float* __restrict__ out) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j < n) out[j] = (float)in[j];
}

}  // namespace

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

int zero_i32(sd_context* ctx, int32_t* p, int64_t n) {
  if (n <= 0) return SD_OK;
  hipLaunchKernelGGL(k_fill_i32, dim3(stride_grid(n)), dim3(256), 0, ctx->stream, p, n);
  SD_HIP(hipGetLastError());
  return SD_OK;
}

void matrix_forget_derived(sd_matrix* m) {
  m->unit_ready = m->centre_ready = m->sq_ready = false;
  for (int t = 0; t < kDenseKinds; ++t) m->dense_kpad[t] = 0;
}

__global__ void k_rows16(const int32_t* __restrict__ in, int64_t n, uint16_t* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = (uint16_t)in[e];
}

int prep_unit(sd_context* ctx, sd_matrix* m) {
  if (m->unit_ready) return SD_OK;
  SD_TRY(m->unit.ensure(sizeof(double) * (size_t)(m->nnz > 0 ? m->nnz : 1)));
  SD_TRY(m->scale.ensure(sizeof(double) * (size_t)(m->ncol + 1)));
  DevBuf& sq = m->unit_ssq;
  SD_TRY(sq.ensure(sizeof(double) * (size_t)(m->ncol + 1)));
  const int64_t nb = (m->ncol + 255) / 256;
  hipLaunchKernelGGL(k_unit_scale, dim3((unsigned)nb), dim3(256), 0, ctx->stream, m->p.as<int64_t>(),
                     m->x.as<double>(), m->ncol, m->scale.as<double>(), sq.as<double>());
  const int64_t nw = (m->ncol + 3) / 4;
  hipLaunchKernelGGL(k_unit_vals, dim3((unsigned)nw), dim3(256), 0, ctx->stream, m->p.as<int64_t>(),
                     m->x.as<double>(), m->ncol, m->scale.as<double>(), sq.as<double>(),
                     m->unit.as<double>());
  if (matrix_rows16(m)) {
    SD_TRY(m->rows16.ensure(sizeof(uint16_t) * (size_t)m->nnz + 16));
    hipLaunchKernelGGL(k_rows16, dim3(stride_grid(m->nnz)), dim3(256), 0, ctx->stream,
                       m->i.as<int32_t>(), m->nnz, m->rows16.as<uint16_t>());
  }
  SD_HIP(hipGetLastError());
  SD_TRY(sync_stream(ctx));
  m->unit_ready = true;
  return SD_OK;
}

int prep_centre(sd_context* ctx, sd_matrix* m) {
  if (m->centre_ready) return SD_OK;
  SD_TRY(m->mean.ensure(sizeof(double) * (size_t)(m->ncol + 1)));
  SD_TRY(m->css.ensure(sizeof(double) * (size_t)(m->ncol + 1)));
  SD_TRY(m->cmf.ensure(sizeof(float) * (size_t)(m->ncol + 1)));
  SD_TRY(m->cif.ensure(sizeof(float) * (size_t)(m->ncol + 1)));
  const int64_t nb = (m->ncol + 255) / 256;
  hipLaunchKernelGGL(k_centre, dim3((unsigned)nb), dim3(256), 0, ctx->stream, m->p.as<int64_t>(),
                     m->x.as<double>(), m->ncol, m->nrow, m->mean.as<double>(), m->css.as<double>());
  hipLaunchKernelGGL(k_centre_f32, dim3((unsigned)nb), dim3(256), 0, ctx->stream,
                     m->mean.as<double>(), m->css.as<double>(), m->ncol, m->nrow,
                     m->cmf.as<float>(), m->cif.as<float>());
  SD_HIP(hipGetLastError());
  SD_TRY(sync_stream(ctx));
  m->centre_ready = true;
  return SD_OK;
}

int prep_sq(sd_context* ctx, sd_matrix* m) {
  if (m->sq_ready) return SD_OK;
  SD_TRY(m->sqn.ensure(sizeof(double) * (size_t)(m->ncol + 1)));
  const int64_t nb = (m->ncol + 255) / 256;
  SD_TRY(m->sqf.ensure(sizeof(float) * (size_t)(m->ncol + 1)));
  hipLaunchKernelGGL(k_sqnorm, dim3((unsigned)nb), dim3(256), 0, ctx->stream, m->p.as<int64_t>(),
                     m->x.as<double>(), m->ncol, m->sqn.as<double>());
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)nb), dim3(256), 0, ctx->stream, m->sqn.as<double>(),
                     m->ncol, m->sqf.as<float>());
  SD_HIP(hipGetLastError());
  SD_TRY(sync_stream(ctx));
  m->sq_ready = true;
  return SD_OK;
}

int prep_dense(sd_context* ctx, sd_matrix* m, int kind, int64_t kpad) {
  if (kind < 0 || kind >= kDenseKinds) return set_error(SD_EINVAL, "internal: dense kind %d", kind);
  if (m->dense_kpad[kind] == kpad && m->dense[kind].ptr) return SD_OK;
  if (kpad % dense_row_align(kind) != 0)
    return set_error(SD_EINVAL, "internal: kpad %lld is not aligned for dense kind %d",
                     (long long)kpad, kind);
  const int64_t row_bytes = dense_row_bytes(kind, kpad);
  m->dense_kpad[kind] = 0;
  SD_TRY(m->dense[kind].ensure((size_t)m->ncol * (size_t)row_bytes + 256));
  uint8_t* dst = m->dense[kind].as<uint8_t>();
  const double* vals = m->x.as<double>();
  if (kind == kDenseBf16Unit || kind == kDenseF32Unit) {
    SD_TRY(prep_unit(ctx, m));
    vals = m->unit.as<double>();
  }
  // rows per chunk: 64 KB of LDS, a multiple of 128 rows (whole 16-byte
  // stores, and whole 32-bit words of eight FP4 rows)
  const int64_t chunk_bytes = 64 * 1024;
  const int64_t chunk = kind == kDenseF4Pattern    ? chunk_bytes * 2
                        : kind == kDenseBf16Split ? chunk_bytes / 4  // hi and lo in LDS
                                                  : chunk_bytes / (row_bytes / kpad);
  const int64_t nchunks = (kpad + chunk - 1) / chunk;
  const int64_t grid = m->ncol * nchunks;
  const size_t lds = kind == kDenseBf16Split ? (size_t)(kpad < chunk ? kpad : chunk) * 4
                                             : (size_t)dense_row_bytes(kind, kpad < chunk ? kpad : chunk);
  if (grid * 256 >= ((int64_t)1 << 32))
    return set_error(SD_EUNSUPPORTED, "densify grid too large (%lld workgroups)", (long long)grid);
  if (grid > 0) {
    switch (kind) {
      case kDenseF4Pattern:
        hipLaunchKernelGGL((k_densify<kDenseF4Pattern, 0>), dim3((unsigned)grid), dim3(256), lds,
                           ctx->stream, m->p.as<int64_t>(), m->i.as<int32_t>(), vals, m->ncol, kpad,
                           chunk, dst);
        break;
      case kDenseBf16Unit:
        hipLaunchKernelGGL((k_densify<kDenseBf16Unit, 2>), dim3((unsigned)grid), dim3(256), lds,
                           ctx->stream, m->p.as<int64_t>(), m->i.as<int32_t>(), vals, m->ncol, kpad,
                           chunk, dst);
        break;
      case kDenseBf16Raw:
        hipLaunchKernelGGL((k_densify<kDenseBf16Raw, 2>), dim3((unsigned)grid), dim3(256), lds,
                           ctx->stream, m->p.as<int64_t>(), m->i.as<int32_t>(), vals, m->ncol, kpad,
                           chunk, dst);
        break;
      case kDenseBf16Split:
        hipLaunchKernelGGL(k_densify_split, dim3((unsigned)grid), dim3(256), lds, ctx->stream,
                           m->p.as<int64_t>(), m->i.as<int32_t>(), vals, m->ncol, kpad, chunk, dst);
        break;
      case kDenseF32Unit:
        hipLaunchKernelGGL((k_densify<kDenseF32Unit, 4>), dim3((unsigned)grid), dim3(256), lds,
                           ctx->stream, m->p.as<int64_t>(), m->i.as<int32_t>(), vals, m->ncol, kpad,
                           chunk, dst);
        break;
      case kDenseF32Pattern:
        hipLaunchKernelGGL((k_densify<kDenseF32Pattern, 4>), dim3((unsigned)grid), dim3(256), lds,
                           ctx->stream, m->p.as<int64_t>(), m->i.as<int32_t>(), vals, m->ncol, kpad,
                           chunk, dst);
        break;
      default:
        hipLaunchKernelGGL((k_densify<kDenseF32Raw, 4>), dim3((unsigned)grid), dim3(256), lds,
                           ctx->stream, m->p.as<int64_t>(), m->i.as<int32_t>(), vals, m->ncol, kpad,
                           chunk, dst);
        break;
    }
    SD_HIP(hipGetLastError());
  }
  SD_TRY(sync_stream(ctx));
  m->dense_kpad[kind] = kpad;
  return SD_OK;
}

}  // namespace sd
