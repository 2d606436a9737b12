// topk_block.h -- exact workgroup-level top-k over a stream of candidates.
//
// Order: (value, index) with the lower index winning ties, values compared
// as doubles (so -0.0 ties with +0.0), non-finite values never admitted --
// the semantics of the reference's bounded heap (src/topK.cpp:42-49, 116-131).
//
// The kept list lives in LDS sorted best first.  Candidates arrive 256 at a
// time (one per thread); each is compared with the current k-th entry and,
// if better, appended after the list.  When anything was appended, list and
// appendix are sorted together with a bitonic network and truncated to k.
// The admission test only ever loosens relative to an exact heap, and the
// sort is over a strict total order, so the result is exact.
#pragma once

#include "sd_internal.h"

namespace sd {

template <int CAP>
struct BlockTopKStorage {
  uint64_t key[CAP];  // smaller = better
  double val[CAP];
  uint32_t idx[CAP];
  int cnt;
  int nbuf;
};

__device__ inline uint64_t order_key(double v, int decreasing) {
  if (v == 0.0) v = 0.0;  // -0.0 compares equal to +0.0
  uint64_t b = (uint64_t)__double_as_longlong(v);
  b ^= (b >> 63) ? ~0ULL : 0x8000000000000000ULL;  // ascending in v
  return decreasing ? ~b : b;
}

template <int CAP>
struct BlockTopK {
  BlockTopKStorage<CAP>& s;
  int k;
  int dec;

  __device__ BlockTopK(BlockTopKStorage<CAP>& st, int k_, int decreasing)
      : s(st), k(k_), dec(decreasing) {
    if (threadIdx.x == 0) {
      s.cnt = 0;
      s.nbuf = 0;
    }
    __syncthreads();
  }

  __device__ static bool less(uint64_t ka, uint32_t ia, uint64_t kb, uint32_t ib) {
    return ka < kb || (ka == kb && ia < ib);
  }

  __device__ void sort_prefix(int n) {
    int P = 2;
    while (P < n) P <<= 1;
    for (int t = n + threadIdx.x; t < P; t += blockDim.x) {
      s.key[t] = ~0ULL;
      s.idx[t] = 0xFFFFFFFFu;
    }
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int t = threadIdx.x; t < P / 2; t += blockDim.x) {
          const int lo = 2 * stride * (t / stride) + (t % stride);
          const int hi = lo + stride;
          const bool up = ((lo & size) == 0);
          const bool gt = less(s.key[hi], s.idx[hi], s.key[lo], s.idx[lo]);
          if (gt == up) {
            const uint64_t tk = s.key[lo];
            s.key[lo] = s.key[hi];
            s.key[hi] = tk;
            const uint32_t ti = s.idx[lo];
            s.idx[lo] = s.idx[hi];
            s.idx[hi] = ti;
            const double tv = s.val[lo];
            s.val[lo] = s.val[hi];
            s.val[hi] = tv;
          }
        }
        __syncthreads();
      }
    }
  }

  // get(t, v, id) -> bool: candidate t (0 <= t < n); false = skip.
  template <class F>
  __device__ void run(int64_t n, F get) {
    for (int64_t base = 0; base < n; base += blockDim.x) {
      const int64_t t = base + threadIdx.x;
      double v = 0.0;
      int64_t id = 0;
      bool ok = false;
      if (t < n) ok = get(t, v, id) && isfinite(v);
      const int cnt = s.cnt;
      uint64_t key = 0;
      if (ok) {
        key = order_key(v, dec);
        if (cnt >= k) ok = less(key, (uint32_t)id, s.key[k - 1], s.idx[k - 1]);
      }
      if (ok) {
        const int slot = atomicAdd(&s.nbuf, 1);
        s.key[cnt + slot] = key;
        s.val[cnt + slot] = v;
        s.idx[cnt + slot] = (uint32_t)id;
      }
      __syncthreads();
      const int nb = s.nbuf;
      if (nb > 0) {
        sort_prefix(cnt + nb);
        if (threadIdx.x == 0) {
          s.cnt = (cnt + nb < k) ? cnt + nb : k;
          s.nbuf = 0;
        }
      }
      __syncthreads();
    }
  }

  // Best first into out[t * ld] for t < k; index + base, NA padding.
  __device__ void write(int32_t* out_idx, double* out_val, int64_t ld, int base) {
    const int cnt = s.cnt;
    for (int t = threadIdx.x; t < k; t += blockDim.x) {
      if (t < cnt) {
        out_idx[(int64_t)t * ld] = (int32_t)s.idx[t] + base;
        out_val[(int64_t)t * ld] = s.val[t];
      } else {
        out_idx[(int64_t)t * ld] = kNaInt;
        out_val[(int64_t)t * ld] = __longlong_as_double((long long)kNaRealBits);
      }
    }
  }
};

}  // namespace sd
