// k_slab.hip -- the slab kernels: densified reference x query tiles on MFMA,
// so the n_ref x n_query slab (the reference's dense block,
// R/sparseKNN.R:256-258) never exists in memory.
//   k_fused        list mode: each lane keeps its query's L best candidates
//                  sorted in registers over a range of reference tiles (small
//                  searches, the threshold pre-pass, the wider second pass).
//   k_filter_wide  filtered full pass: every value at or above its query's
//                  threshold is appended to a per-query candidate buffer.
// Plus the kernels that finish a search: merging per-split lists, selecting
// from the candidate buffers, the FP64 re-rank with its certificate, and
// writing FNN-layout results.
//
// k_fused: workgroup = WM x 4 waves.  Wave (wm, wn) computes 128 reference
// columns x 64 query columns with 32x32 MFMA blocks (4 along refs, 2 along
// queries).  After each reference tile the two lanes holding the same query
// exchange values (lane l owns query block 0, lane l+32 query block 1), so
// one lane owns one query and keeps its L best candidates sorted in registers.
#include <hip/hip_bf16.h>

#include <cfloat>

#include "pair_math.h"
#include "sd_internal.h"
#include "topk_block.h"

namespace sd {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace {

constexpr int kTileRefs = 128;  // reference columns per wave row
constexpr int kRowBytes = 128;  // K bytes per row per chunk
// Word 3 of a raw buffer resource on gfx9-family parts: 32-bit data format,
// no swizzle, range checking against num_records in bytes.
constexpr int kRsrcWord3 = 0x00020000;

// ---- MFMA operand types ------------------------------------------------------
// mma() consumes one 16-byte K segment of each operand per lane.  Within one
// MFMA, lane (m, h) supplies K elements of its own choosing; A and B lanes
// with the same h always supply the same logical segment, so the product is
// the full dot product whatever the hardware's K order inside a lane.
struct OpF4 {  // 0/1 patterns as FP4 (e2m1) nibbles: 1.0 products, exact FP32 counts
  using Acc = v16f;  // counts stay exact below 2^24 (rows are far fewer)
  __device__ static void mma(Acc& c, const uint4& a, const uint4& b) {
    // 32 K elements per lane; only the low four dwords of an FP4 operand
    // are read.  Both block scales are 127 (E8M0 for 2^0).
    const v8i av = {(int)a.x, (int)a.y, (int)a.z, (int)a.w, 0, 0, 0, 0};
    const v8i bv = {(int)b.x, (int)b.y, (int)b.z, (int)b.w, 0, 0, 0, 0};
    c = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, c, 4, 4, 0, 127, 0, 127);
  }
};

struct OpBf16 {  // BF16 values, FP32 accumulation
  using Acc = v16f;
  __device__ static void mma(Acc& c, const uint4& a, const uint4& b) {
    v8bf av, bv;
    __builtin_memcpy(&av, &a, 16);
    __builtin_memcpy(&bv, &b, 16);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, c, 0, 0, 0);
  }
};

struct OpF32 {  // plain FP32 values and products
  using Acc = v16f;
  __device__ static void mma(Acc& c, const uint4& a, const uint4& b) {
    float av[4], bv[4];
    __builtin_memcpy(av, &a, 16);
    __builtin_memcpy(bv, &b, 16);
#pragma unroll
    for (int e = 0; e < 4; ++e) c = __builtin_amdgcn_mfma_f32_32x32x2f32(av[e], bv[e], c, 0, 0, 0);
  }
};

// ---- candidate keys --------------------------------------------------------------
// Jaccard: the exact rational sim = c / d (d > 0; two empty columns are 1/1,
// src/sparseDist.cpp:551-556).  Distance and similarity forms rank by the
// same order of sim (1 - sim is monotone, and distinct rationals with
// denominators below 2^26 stay distinct and ordered in FP64), so ranking by
// the rational is ranking by the reference's double.
struct KeyJ {
  uint32_t c, d;
};
__device__ inline bool better(const KeyJ& a, uint32_t ia, const KeyJ& b, uint32_t ib) {
  const uint64_t l = (uint64_t)a.c * b.d, r = (uint64_t)b.c * a.d;
  return l > r || (l == r && ia < ib);
}
__device__ inline KeyJ worst_key(KeyJ*) { return KeyJ{0u, 1u}; }
__device__ inline double key_to_double(const KeyJ& k) { return (double)k.c / (double)k.d; }

// Approximate similarity (larger is better).
__device__ inline bool better(float a, uint32_t ia, float b, uint32_t ib) {
  return a > b || (a == b && ia < ib);
}
__device__ inline float worst_key(float*) { return -INFINITY; }
__device__ inline double key_to_double(float k) { return (double)k; }

// Euclidean key: minus a lower bound of the squared distance
// D = |q|^2 + |r|^2 - 2 q.r, from FP32 squared norms qa, ra and the slab's
// dot x.  With e = e0 + e1 nnz(q) the dot obeys |x - q.r| <= e S, where
// S = sum |q_i r_i| <= |q||r| <= (qa + ra) / 2; for non-negative data S is
// q.r itself and q.r <= x / (1 - e) (the host folds 1 / (1 - e) into e0, e1).
// So D >= s - err with s = qa + ra - 2x and err = 2e |x| + kEucRound (qa + ra)
// (non-negative data) or (e + kEucRound) (qa + ra) (signed).  kEucRound covers
// the FP32 rounding of qa, ra and of this formula (at most six roundings of
// quantities below 2 (qa + ra), 2^-24 each); kEucAbs covers flushed
// subnormal products.  The per-query weights w1, w2 select the case without
// a branch (a branch per value made the compiler spill); larger keys are
// better, like every key here.
constexpr float kEucRound = 4.76837158203125e-07f;  // 2^-21
constexpr float kEucAbs = 1e-30f;
struct KeyWeights {
  float w1, w2, w3;
};
__device__ inline KeyWeights euc_weights(uint32_t qnnz, float e0, float e1, int nonneg) {
  const float e = e0 + e1 * (float)qnnz;
  return KeyWeights{nonneg ? 2.f * e : 0.f, nonneg ? kEucRound : e + kEucRound, 0.f};
}
__device__ inline float euc_key(const KeyWeights& w, float qa, float ra, float x) {
  const float n2 = qa + ra;
  const float s = n2 - 2.f * x;
  const float err = (w.w1 * fabsf(x) + w.w2 * n2) + kEucAbs;
  return err - s;
}
// Pearson key: an upper bound of the correlation
// r = (q.y - n mq my) / (sq sy) from FP32 aux a = sqrt(n) mean and
// c = 1 / sqrt(centred sum of squares) per column and the slab's raw dot x.
// The dot obeys |x - q.y| <= e S with S = sum |q_i y_i|: S = q.y <= x / (1 - e)
// for non-negative data (folded into e by the host), else
// S <= |q||y| = sq sy g_q g_y with g = sqrt(1 + (a c)^2).  The bound adds
// (e + kPearRound) S c_q c_y + kPearRound (|x| + |t|) c_q c_y to the estimate;
// kPearRound covers the FP32 rounding of the aux values and of this formula
// (at most ten roundings, 2^-24 each, of quantities below (|x| + |t|) c_q c_y
// or S c_q c_y).  Weights: w1 = e + kPearRound on |x| c_q c_y (non-negative
// data) and w2 = (e + kPearRound) g_q on g_y (signed), the other one 0.
constexpr float kPearRound = 9.5367431640625e-07f;  // 2^-20
__device__ inline KeyWeights pear_weights(float qa, float qc, uint32_t qnnz, float e0, float e1,
                                          int nonneg) {
  const float ew = (e0 + e1 * (float)qnnz) + kPearRound;
  const float gq = sqrtf(1.f + (qa * qc) * (qa * qc));
  return KeyWeights{nonneg ? ew : 0.f, nonneg ? 0.f : ew * gq, 0.f};
}
__device__ inline float pear_key(const KeyWeights& w, float qa, float qc, float ra, float rc,
                                 float x) {
  const float t = qa * ra;  // n mean_q mean_y
  const float sc = qc * rc;
  const float r = (x - t) * sc;
  const float gy = sqrtf(1.f + (ra * rc) * (ra * rc));
  return r + (((w.w1 * (fabsf(x) * sc) + w.w2 * gy) + kPearRound * (fabsf(x) + fabsf(t)) * sc) +
              kEucAbs);
}
// The next float above a finite t (t itself for +inf and NaN).
__device__ inline float next_float_up(float t) {
  if (!(t < INFINITY)) return t;
  if (t == 0.f) return __int_as_float(1);
  const int b = __float_as_int(t);
  return __int_as_float(t > 0.f ? b + 1 : b - 1);
}

// Jaccard screen in one FMA: with d = ni + nj - c > 0, c / d >= t exactly
// when c >= t (ni + nj) / (1 + t), so a value passes the screen when
// c >= fma(a, ni, a nj) with a = t / (1 + t), t the screen threshold (a
// little below the exact one, at least 0).  a is lowered by a further 2^-18
// relative: the FP32 quotient, product and FMA round by at most 2^-22
// relative in all, so the screen never rejects a value the exact test keeps.
// Two empty columns (c = ni = nj = 0) pass (J = 1).  An infinite threshold
// (no query) gives a NaN coefficient, which nothing passes.
__device__ inline float jac_screen_coef(float t) {
  const float tt = fmaxf(t, 0.f);
  return (tt / (1.f + tt)) * 0.999996185302734375f;  // 1 - 2^-18
}

// Split layout's query offset: segment 1 <-> 2.  Written as arithmetic, not
// selects, so that the K loop stays one basic block (a branch there makes the
// wait-count pass merge states at the join and wait for every load), and
// read back as a wave-uniform scalar: the buffer load's scalar offset must
// not come from a vector register (the compiler would loop over lanes).
__device__ inline int qseg_offset(int soff, int qseg) {
  const int in1 = (int)(soff >= qseg) & (int)(soff < 2 * qseg);
  const int in2 = (int)(soff >= 2 * qseg);
  return __builtin_amdgcn_readfirstlane(soff + qseg * (in1 - in2));
}

template <class K>
struct Entry {
  K key;
  uint32_t idx;
};

}  // namespace

struct FusedKernelArgs {
  const uint8_t* ref_dense;   // chunk-major (dense_offset) with ref_ld ...
  const uint8_t* qry_dense;   // ... and qry_ld columns per chunk block
  int64_t ref_ld, qry_ld;
  const int32_t* qlist;  // optional: position -> query column
  int64_t nref, nqry, kpad_bytes;
  const int64_t* ref_p;
  const int64_t* qry_p;
  const float* ref_aux;
  const float* qry_aux;
  const float* ref_aux2;
  const float* qry_aux2;
  const uint8_t* ref_deg;
  int self_mode, include_self, method;
  int64_t self_off;     // self mode: query column j is reference column j + self_off
  int splits, qtiles, tiles_per_split;
  int64_t ntiles;       // list mode: virtual tiles (every tile_stride-th reference tile)
  int tile_stride;
  double* cand_key;
  int32_t* cand_idx;
  // filter mode: one (query tile, reference tile) per workgroup
  int rtiles, qgroups;
  const double* thr;    // per query position: keep candidates with key >= thr
  int cap;
  int32_t* cnt;         // per query position: candidates appended
  double* fkey;         // [nqry][cap]
  int32_t* fidx;
  // split layouts: query rows are read with the segments at byte offsets
  // [qseg, 2 qseg) and [2 qseg, 3 qseg) swapped (0: read as stored)
  int qseg;
  float dot_e0, dot_e1;  // euclidean: dot error bound (e0 + e1 nnz(q)) sum |q_i r_i|
  int dot_nonneg;        // both matrices >= 0: that sum is the dot product
  // filter mode, mirror (k_filter_wide with SYM): every value also screened
  // for the reference column as a query of the mirrored search, whose
  // per-query thresholds, counts and buffers (cap entries each) these are,
  // with the query column as its candidate (mdeg: the candidates' degenerate
  // flags).  tri = 1: a self-search, the mirrored search is this one and only
  // tiles with rt >= qt run; tri = 0: a cross search, every tile runs and
  // feeds the reverse search.
  const double* mthr;
  int mcap;
  int32_t* mcnt;
  double* mfkey;
  int32_t* mfidx;
  const uint8_t* mdeg;
  int tri;
};

namespace {

// XCD-aware block mapping: blocks are dispatched round-robin over the 8
// XCDs, so consecutive block ids sit on different L2s.  Give each XCD whole
// groups of (query tile, every split) so that the concurrent workgroups of
// one XCD share their query tiles (B) and reference tiles (A) in its L2.
__device__ inline void map_block(int b, int qtiles, int splits, int& qt, int& sp) {
  const int nblk = qtiles * splits;
  int lin = b;
  if ((nblk & 7) == 0) lin = (b & 7) * (nblk >> 3) + (b >> 3);  // contiguous range per XCD
  qt = lin / splits;
  sp = lin % splits;
}

// Filter-mode rasterisation: tiles are ordered (group of 4 query tiles,
// reference tile, query tile in group) and each XCD takes a contiguous range
// of that order, so the 32 workgroups resident on one XCD cover 4 query tiles
// x 8 reference tiles and share those A and B tiles in the XCD's L2.
__device__ inline bool map_tile(int64_t b, int64_t nblk, int qtiles, int rtiles, int& qt,
                                int64_t& rt) {
  const int64_t lin = (b & 7) * (nblk >> 3) + (b >> 3);
  const int64_t per = 4 * (int64_t)rtiles;
  const int64_t qg = lin / per, rem = lin % per;
  rt = rem / 4;
  qt = (int)(qg * 4 + rem % 4);
  return qt < qtiles;
}

// Symmetric self-search: the query and reference tiles are the same columns,
// so only tiles with rt >= qt are visited, in the same (group of 4 query
// tiles, reference tile, query tile in group) order; group g covers reference
// tiles 4g .. rtiles - 1, 4 (rtiles - 4g) slots, of which the few with
// rt < qt near the diagonal stay empty.  Slots before group g:
// cum(g) = 4 g rtiles - 8 g (g - 1).  The order is dealt to the XCDs in
// chunks of kSymChunk slots, round robin (the grid is a multiple of
// 8 kSymChunk blocks): the XCDs then work side by side on neighbouring
// reference tiles of the same query tiles, so the query tiles come from the
// shared last-level cache; with one contiguous range per XCD (the
// rectangular order's scheme) the triangle leaves each XCD on a different
// stretch of references, and the pass ran 4% slower.
constexpr int kSymChunk = 32;
__device__ inline int64_t sym_slots_before(int64_t g, int64_t rtiles) {
  return 4 * g * rtiles - 8 * g * (g - 1);
}
__device__ inline bool map_tile_sym(int64_t b, int qtiles, int rtiles, int& qt, int64_t& rt) {
  const int64_t k = b >> 3;  // the k-th block dispatched to XCD b % 8
  const int64_t lin = ((k / kSymChunk) * 8 + (b & 7)) * kSymChunk + k % kSymChunk;
  const int64_t groups = (qtiles + 3) / 4;
  if (lin >= sym_slots_before(groups, rtiles)) return false;
  int64_t lo = 0, hi = groups;  // cum(lo) <= lin < cum(hi)
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (sym_slots_before(mid, rtiles) <= lin) lo = mid;
    else hi = mid;
  }
  const int64_t rem = lin - sym_slots_before(lo, rtiles);
  rt = 4 * lo + rem / 4;
  qt = (int)(4 * lo + rem % 4);
  return qt < qtiles && rt >= qt;
}

// Workgroup barrier for LDS only: wait for this wave's LDS operations, then
// s_barrier.  Unlike __syncthreads() it does not wait for outstanding global
// loads, which is what lets direct-to-LDS loads stay in flight across it.
// The builtin wait comes first so that the compiler's own wait accounting
// knows that every LDS and scalar load has completed here.
__device__ inline void lds_barrier() {
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0), other counters untouched
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// s_waitcnt vmcnt(N) with the other counters left alone (gfx9 encoding:
// vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
template <int N>
__device__ inline void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt is six bits");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x0F70);
}

// Key kinds of the FP32-accumulator kernels, fixed at compile time so each
// instantiation carries only its own epilogue: the score itself (cosine),
// minus a squared-distance lower bound (euclidean), a correlation upper
// bound (pearson).  QS: query rows are read with the split layout's
// segments swapped (qseg).
enum KeyKind { kKeyScore = 0, kKeyEuclid = 1, kKeyPearson = 2 };

template <class Op, class K, int L, int WM, int MK, bool QS>
__global__ __launch_bounds__(WM * 256, 1) void k_fused(FusedKernelArgs P) {
  constexpr int WAVES = WM * 4;
  constexpr int NT = WAVES * 64;
  constexpr int TM = kTileRefs * WM;  // refs per workgroup tile
  constexpr int TN = 256;             // queries per workgroup
  // K streams through a ring of three LDS stages filled by direct
  // global->LDS loads (global_load_lds_dwordx4): each wave instruction writes
  // 64 lanes x 16 B = 1 KB contiguously, i.e. 8 rows x 128 B.  Rows are not
  // padded; the 16-byte slot of a segment is XOR-swizzled by (row >> 1) & 7
  // so that 16 consecutive lanes reading one row each hit all 64 banks once.
  // Each stage is its own LDS object and the loop is unrolled by three, so
  // the compiler sees that the loads in flight never alias the stage being
  // read and waits only for the chunk it needs.  The epilogue's lane
  // exchange (NT lanes x 33 words) reuses stage 0 once the ring is drained.
  static_assert(WM == 1, "the staging layout assumes one wave row");
  constexpr int A_BYTES = TM * kRowBytes;  // 16 KB
  constexpr int B_BYTES = TN * kRowBytes;  // 32 KB
  constexpr int STAGE_BYTES = A_BYTES + B_BYTES;
  static_assert(NT * 33 * 4 <= STAGE_BYTES, "exchange area must fit in one stage");
  __shared__ __align__(16) uint8_t st0[STAGE_BYTES];
  __shared__ __align__(16) uint8_t st1[STAGE_BYTES];
  __shared__ __align__(16) uint8_t st2[STAGE_BYTES];
  uint32_t* xch = reinterpret_cast<uint32_t*>(st0);
  __shared__ __align__(16) int32_t s_rnnz[TM];  // read four at a time by the Jaccard marks
  __shared__ float s_raux[TM];
  __shared__ float s_raux2[TM];
  __shared__ uint8_t s_rdeg[TM];

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = wave / 4, wn = wave % 4;
  const int h = lane >> 5, n = lane & 31;
  int qt, sp;
  map_block(blockIdx.x, P.qtiles, P.splits, qt, sp);
  const int64_t q_base = (int64_t)qt * TN;
  // position of the query this lane owns, and its column in the query matrix
  const int64_t my_pos = q_base + wn * 64 + h * 32 + n;
  const bool q_ok = my_pos < P.nqry;
  const int64_t my_q = q_ok ? (P.qlist ? (int64_t)P.qlist[my_pos] : my_pos) : -1;
  uint32_t nj = 0;
  float qaux = 0.f, qaux2 = 0.f;
  if (q_ok) {
    nj = (uint32_t)(P.qry_p[my_q + 1] - P.qry_p[my_q]);
    if (P.qry_aux) qaux = P.qry_aux[my_q];
    if (P.qry_aux2) qaux2 = P.qry_aux2[my_q];
  }
  const KeyWeights kw = MK == kKeyPearson
                            ? pear_weights(qaux, qaux2, nj, P.dot_e0, P.dot_e1, P.dot_nonneg)
                            : euc_weights(nj, P.dot_e0, P.dot_e1, P.dot_nonneg);

  Entry<K> list[L];
#pragma unroll
  for (int t = 0; t < L; ++t) {
    list[t].key = worst_key((K*)nullptr);
    list[t].idx = 0xFFFFFFFFu;
  }

  const int64_t t0 = (int64_t)sp * P.tiles_per_split;
  const int64_t t1 = (t0 + P.tiles_per_split < P.ntiles) ? t0 + P.tiles_per_split : P.ntiles;
  const int64_t nchunks = P.kpad_bytes / kRowBytes;

  // Direct-load sources.  Wave w issues A instructions 4w..4w+3 (of 16) and
  // B instructions 8w..8w+7 (of 32); lane l of instruction i carries row
  // 8i + l/8 and physical slot l%8, i.e. logical segment
  // (l%8) ^ ((row >> 1) & 7) = (l%8) ^ ((4i + l/16) & 7).
  constexpr int A_INST = A_BYTES / 1024 / WAVES;  // 4
  constexpr int B_INST = B_BYTES / 1024 / WAVES;  // 8
  auto seg_of = [&](int i) { return (lane & 7) ^ ((4 * i + (lane >> 4)) & 7); };
  // The loads go through buffer resources (buffer_load_dwordx4 ... lds) rather
  // than global_load_lds: the compiler treats an LDS-writing FLAT load as
  // pending on both counters and would then round every LDS wait in the loop
  // down to lgkmcnt(0), serialising the fragment reads with the MFMAs.  A
  // resource covers one chunk block of a tile (rows 128 bytes apart in the
  // chunk-major layout); rows past the end are clamped to the last row (they
  // feed only masked or unused lanes).  Rows are addressed by query
  // position: the host gathers listed queries first.
  const int64_t b_rows = (P.nqry - q_base < TN) ? P.nqry - q_base : TN;
  uint32_t vb[B_INST];
#pragma unroll
  for (int u = 0; u < B_INST; ++u) {
    const int i = wave * B_INST + u;
    int row = 8 * i + (lane >> 3);
    if (row >= b_rows) row = (int)b_rows - 1;
    vb[u] = (uint32_t)(row * kDenseChunk) + seg_of(i) * 16;
  }
  typedef __attribute__((address_space(3))) void lds_void;

  for (int64_t vt = t0; vt < t1; ++vt) {
    const int64_t tile = vt * P.tile_stride;
    const int64_t r_base = tile * TM;
    for (int t = tid; t < TM; t += NT) {
      const int64_t r = r_base + t;
      if (r < P.nref) {
        s_rnnz[t] = (int32_t)(P.ref_p[r + 1] - P.ref_p[r]);
        s_raux[t] = P.ref_aux ? P.ref_aux[r] : 0.f;
        s_raux2[t] = P.ref_aux2 ? P.ref_aux2[r] : 0.f;
        s_rdeg[t] = P.ref_deg ? P.ref_deg[r] : 0;
      } else {
        s_rnnz[t] = 0;
        s_raux[t] = 0.f;
        s_raux2[t] = 0.f;
        s_rdeg[t] = 1;
      }
    }

    typename Op::Acc acc[4][2];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[b][j][r] = 0;

    const int64_t a_rows = (P.nref - r_base < TM) ? P.nref - r_base : TM;
    uint32_t va[A_INST];
#pragma unroll
    for (int u = 0; u < A_INST; ++u) {
      const int i = wave * A_INST + u;
      int row = 8 * i + (lane >> 3);
      if (row >= a_rows) row = (int)a_rows - 1;  // clamped rows are masked by s_rdeg
      va[u] = (uint32_t)(row * kDenseChunk) + seg_of(i) * 16;
    }
    // chunk c of both tiles: one block each of the chunk-major layouts
    auto issue = [&](uint8_t* st, int64_t c) {
      const int kb = (int)(c * kRowBytes);
      const int kb_b = QS ? qseg_offset(kb, P.qseg) : kb;
      const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(P.ref_dense + dense_offset(r_base, kb, P.ref_ld)), 0,
          (int)(a_rows * kDenseChunk), kRsrcWord3);
      const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(P.qry_dense + dense_offset(q_base, kb_b, P.qry_ld)), 0,
          (int)(b_rows * kDenseChunk), kRsrcWord3);
#pragma unroll
      for (int u = 0; u < A_INST; ++u)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            ra, (lds_void*)(st + (wave * A_INST + u) * 1024), 16, (int)va[u], 0, 0, 0);
#pragma unroll
      for (int u = 0; u < B_INST; ++u)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rb, (lds_void*)(st + A_BYTES + (wave * B_INST + u) * 1024), 16, (int)vb[u], 0, 0, 0);
    };
    const int swz = (n >> 1) & 7;
    // Operand fragments of one MFMA step (16 K bytes per lane): lanes with h
    // supply logical segment 2t + h.  Two register sets alternate so that the
    // LDS reads of step t + 1 are in flight while step t's MFMAs run; with one
    // wave per SIMD nothing else hides their latency.
    struct Frag {
      uint4 a[4], b[2];
    };
    auto load_frag = [&](Frag& f, const uint8_t* st, int t) {
      const int slot = ((2 * t + h) ^ swz) * 16;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        f.a[b] = *reinterpret_cast<const uint4*>(st + (b * 32 + n) * kRowBytes + slot);
#pragma unroll
      for (int j = 0; j < 2; ++j)
        f.b[j] = *reinterpret_cast<const uint4*>(st + A_BYTES + (wn * 64 + j * 32 + n) * kRowBytes +
                                                 slot);
    };
    auto mma_frag = [&](const Frag& f) {
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int j = 0; j < 2; ++j) Op::mma(acc[b][j], f.a[b], f.b[j]);
    };
    Frag f0, f1;

    lds_barrier();  // the previous tile's epilogue is done with the stages
    // kpad is a positive multiple of 128 rows, so there is at least one
    // chunk.  Refills past the last chunk reload the last chunk into a stage
    // nobody reads again: every step then issues the same loads and waits for
    // the same count, and no LDS read sits behind a branch (at a join the
    // compiler would wait for all loads in flight before reading LDS).
    const int64_t last = nchunks - 1;
    issue(st0, 0);
    issue(st1, last < 1 ? last : 1);
    issue(st2, last < 2 ? last : 2);
    wait_vm<2 * (A_INST + B_INST)>();
    lds_barrier();
    load_frag(f0, st0, 0);
    // One chunk c in stage CUR (step 0's fragments already requested): steps
    // 0-2 alternate fragment sets; before step 3's MFMAs, wait until this
    // wave's loads of chunk c + 1 have landed (loads retire in order, so the
    // loads of chunk c + 2 may still be outstanding), barrier -- after which
    // no wave reads CUR any more and every wave's part of chunk c + 1 is
    // visible -- refill CUR with chunk c + 3 and request chunk c + 1's first
    // fragments, all while step 3's MFMAs keep the matrix core busy.
    // Scheduling barriers keep the compiler from sinking the reads next to
    // the MFMAs that consume them.
#define SD_STEPS_0_2(CUR)                                                     \
    load_frag(f1, CUR, 1);                                                    \
    __builtin_amdgcn_sched_barrier(0);                                        \
    mma_frag(f0);                                                             \
    __builtin_amdgcn_sched_barrier(0);                                        \
    load_frag(f0, CUR, 2);                                                    \
    __builtin_amdgcn_sched_barrier(0);                                        \
    mma_frag(f1);                                                             \
    __builtin_amdgcn_sched_barrier(0);                                        \
    load_frag(f1, CUR, 3);                                                    \
    __builtin_amdgcn_sched_barrier(0);                                        \
    mma_frag(f0);                                                             \
    __builtin_amdgcn_sched_barrier(0);
#define SD_CHUNK(CUR, NXT, C)                                                 \
    {                                                                         \
      SD_STEPS_0_2(CUR)                                                       \
      wait_vm<A_INST + B_INST>();                                             \
      lds_barrier();                                                          \
      issue(CUR, (C) + 3 < last ? (C) + 3 : last);                            \
      load_frag(f0, NXT, 0);                                                  \
      __builtin_amdgcn_sched_barrier(0);                                      \
      mma_frag(f1);                                                           \
      __builtin_amdgcn_sched_barrier(0);                                      \
    }
#define SD_LAST_CHUNK(CUR)                                                    \
    {                                                                         \
      SD_STEPS_0_2(CUR)                                                       \
      mma_frag(f1);                                                           \
    }
    int64_t c = 0;
    for (; c + 3 < nchunks; c += 3) {
      SD_CHUNK(st0, st1, c)
      SD_CHUNK(st1, st2, c + 1)
      SD_CHUNK(st2, st0, c + 2)
    }
    // the last one to three chunks (chunk c sits in st0)
    if (nchunks - c == 3) {
      SD_CHUNK(st0, st1, c)
      SD_CHUNK(st1, st2, c + 1)
      SD_LAST_CHUNK(st2)
    } else if (nchunks - c == 2) {
      SD_CHUNK(st0, st1, c)
      SD_LAST_CHUNK(st1)
    } else {
      SD_LAST_CHUNK(st0)
    }
#undef SD_LAST_CHUNK
#undef SD_CHUNK
#undef SD_STEPS_0_2
    wait_vm<0>();  // the trailing refills land before stage 0 is reused
    lds_barrier();  // every wave is done reading the stages

    // ---- epilogue: this lane owns query my_q and must see its 128 values.
    // acc[b][j][r] belongs to ref 32b + 8(r/4) + 4h + (r%4), query block j.
    // Sub-phases 0-1 scan the lane's own block-h values (reference blocks
    // 0-1, then 2-3), sub-phases 2-3 the block-h values held by lane l ^ 32
    // (the other four rows of every group of eight).  Values pass through
    // LDS (rows of 33 words against bank conflicts) so the scan is a loop.
    const int rloc_base = wm * kTileRefs;
    // Plain float keys (cosine) and Jaccard: the lane that stores a row's 32
    // values also marks those at or above the owner's current L-th key in the
    // row's spare word, and the owner visits the marked values only.  The
    // L-th key only rises during the scan, so every value the list takes is
    // marked.  Jaccard marks use the one-FMA screen c >= a n_r + b with a and
    // b from the owner's L-th key (jac_screen_coef, which never drops a count
    // whose exact key reaches the threshold).  A tile that may hold a query's
    // own column marks everything (that value's key is restored above its
    // slab value).
    constexpr bool kListScreen = MK == kKeyScore;
    auto diag = [&](int64_t off) { return r_base < q_base + off + TN && q_base + off < r_base + TM; };
    const bool mark_all = P.self_mode && (P.qlist != nullptr || diag(0) || diag(P.self_off));
#pragma unroll
    for (int sub = 0; sub < 4; ++sub) {
      const int phase = sub >> 1, b0 = (sub & 1) * 2;
      // the row owner's (this lane, then lane l ^ 32) L-th key as a float
      // threshold, or its Jaccard screen coefficients
      float thr_w = 0.f, ja_w = 0.f, jb_w = 0.f;
      if constexpr (kListScreen && sizeof(K) == 8) {
        const KeyJ lk = list[L - 1].key;  // d >= 1
        const float f = (float)((double)lk.c / (double)lk.d);
        const float a_own = jac_screen_coef(f - fabsf(f) * 9.5367431640625e-07f - 1e-30f);
        const float b_own = a_own * (float)nj;
        const float a_oth = __shfl_xor(a_own, 32), b_oth = __shfl_xor(b_own, 32);
        ja_w = (phase == 0) ? a_own : a_oth;
        jb_w = (phase == 0) ? b_own : b_oth;
      } else if constexpr (kListScreen) {
        const float own = (float)list[L - 1].key;
        const float other = __shfl_xor(own, 32);
        thr_w = (phase == 0) ? own : other;
      }
      uint32_t marks = 0u;
#pragma unroll
      for (int bb = 0; bb < 2; ++bb) {
        // stored entries of the references of this lane's values: the row
        // owner's values of a block are the references of this lane's layout
        // (32 b + 8 (r / 4) + 4 h + r % 4) in either phase
        int32_t rn[16];
        if constexpr (kListScreen && sizeof(K) == 8) {
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            const int4 n4 =
                *reinterpret_cast<const int4*>(&s_rnnz[rloc_base + (b0 + bb) * 32 + 8 * g + 4 * h]);
            rn[4 * g] = n4.x;
            rn[4 * g + 1] = n4.y;
            rn[4 * g + 2] = n4.z;
            rn[4 * g + 3] = n4.w;
          }
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const auto v0 = acc[b0 + bb][0][r], v1 = acc[b0 + bb][1][r];
          const auto x = (phase == 0) ? (h ? v1 : v0) : (h ? v0 : v1);
          uint32_t bits;
          if constexpr (sizeof(K) == 8)
            bits = (uint32_t)x;  // exact integer count (int or float accumulator)
          else
            __builtin_memcpy(&bits, &x, 4);
          xch[((phase == 0) ? tid : (tid ^ 32)) * 33 + bb * 16 + r] = bits;
          if constexpr (kListScreen && sizeof(K) == 8)
            marks |= ((float)x >= fmaf(ja_w, (float)rn[r], jb_w)) ? 1u << (bb * 16 + r) : 0u;
          else if constexpr (kListScreen)
            marks |= ((float)x >= thr_w) ? 1u << (bb * 16 + r) : 0u;
        }
      }
      if constexpr (kListScreen)
        xch[((phase == 0) ? tid : (tid ^ 32)) * 33 + 32] = mark_all ? 0xFFFFFFFFu : marks;
      lds_barrier();
      const int hh = (phase == 0) ? h : 1 - h;
      // value v of the row: the exact key and the list insertion
      auto visit = [&](int v) {
        const int b = b0 + (v >> 4), rr = v & 15;
        const int rloc = rloc_base + b * 32 + 8 * (rr >> 2) + 4 * hh + (rr & 3);
        const int64_t rg = r_base + rloc;
        if (s_rdeg[rloc] || (P.self_mode && rg == my_q + P.self_off && !P.include_self)) return;
        const uint32_t bits = xch[tid * 33 + v];
        K key;
        if constexpr (sizeof(K) == 8) {
          const uint32_t c = bits;
          const uint32_t d = (uint32_t)s_rnnz[rloc] + nj - c;
          key = (d == 0) ? KeyJ{1u, 1u} : KeyJ{c, d};
        } else {
          float sv;
          __builtin_memcpy(&sv, &bits, 4);
          if constexpr (MK == kKeyEuclid) {
            key = euc_key(kw, qaux, s_raux[rloc], sv);
            if (P.self_mode && rg == my_q) key = FLT_MAX;  // distance 0
          } else if constexpr (MK == kKeyPearson) {
            key = pear_key(kw, qaux, qaux2, s_raux[rloc], s_raux2[rloc], sv);
            if (P.self_mode && rg == my_q) key = FLT_MAX;  // restored self correlation
          } else {
            key = sv;
            if (P.self_mode && rg == my_q) key = 2.0f;  // restored self similarity
          }
        }
        if (better(key, (uint32_t)rg, list[L - 1].key, list[L - 1].idx)) {
          Entry<K> carry{key, (uint32_t)rg};
#pragma unroll
          for (int p = 0; p < L; ++p) {
            if (better(carry.key, carry.idx, list[p].key, list[p].idx)) {
              const Entry<K> tmp = list[p];
              list[p] = carry;
              carry = tmp;
            }
          }
        }
      };
      if (q_ok) {
        if constexpr (kListScreen) {
          uint32_t m = xch[tid * 33 + 32];
          while (m) {
            const int v = __builtin_ctz(m);
            m &= m - 1u;
            visit(v);
          }
        } else {
#pragma unroll 1
          for (int v = 0; v < 32; ++v) visit(v);
        }
      }
      lds_barrier();
    }
  }

  if (q_ok) {
    const int64_t slot = ((int64_t)sp * WM + wm) * P.nqry + my_pos;
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const bool valid = list[t].idx != 0xFFFFFFFFu;
      P.cand_key[slot * L + t] = valid ? key_to_double(list[t].key) : 0.0;
      P.cand_idx[slot * L + t] = valid ? (int32_t)list[t].idx : -1;
    }
  }
}

// ---- filtered full pass: 256 x 256 tiles, eight waves ------------------------
// One workgroup per (reference tile of 256, query tile of 256); every value
// that reaches its query's threshold is appended to that query's candidate
// buffer.  Wave (wm, wn) computes references 128 wm .. +128 x queries
// 64 wn .. +64, as in k_fused, but a workgroup of eight waves (two per SIMD)
// shares each staged byte among 1.5x the MFMA work of a 128 x 256 tile, and
// the second wave on each SIMD covers the first one's barrier and LDS waits.
// K streams through four LDS stages of 64-byte row chunks (A and B 16 KB
// each); a stage is refilled with the chunk four ahead as soon as every wave
// has read it, so up to three chunks (96 KB) are in flight per CU.
constexpr int kWideTile = 256;  // references and queries per workgroup

// Staging layouts of the filtered pass by K bytes per row per chunk (RB):
// RB = 64 runs four 32 KB stages, RB = 128 two 64 KB stages.  A wave
// instruction writes 1 KB = (1024 / RB) rows x RB bytes; the 16-byte slot of
// logical segment s in row R is s ^ ((R / (256 / RB)) % (RB / 16)), so 16
// lanes reading one segment of 16 consecutive rows hit all 64 banks once.
template <int RB>
struct WideLayout {
  static_assert(RB == 64 || RB == 128, "64- or 128-byte chunks");
  static constexpr int kStages = RB == 64 ? 4 : 2;
  static constexpr int kSlots = RB / 16;          // 16-byte segments per row chunk
  static constexpr int kRowsPerInst = 64 / kSlots;
  static constexpr int kSteps = RB / 32;          // MFMA steps per chunk
  static constexpr int kRowShift = RB == 64 ? 2 : 1;  // rows per 256 bytes = 1 << shift
  __device__ static int swz(int row) { return (row >> kRowShift) & (kSlots - 1); }
};

// SYM: mirrored screening.  A self-search over all columns at once
// (P.tri = 1) runs only tiles with rt >= qt, and a tile with rt > qt also
// screens every value for the reference column as the query and the query
// column as its candidate.  A cross search (P.tri = 0) runs every tile and
// screens every value for the reverse search as well (the reference column
// as its query, the query column as its candidate).
template <class Op, class K, int RB, int MK, bool QS, bool SYM>
__global__ __launch_bounds__(512, 2) void k_filter_wide(FusedKernelArgs P) {
  using Lay = WideLayout<RB>;
  constexpr int NT = 512;
  constexpr int NS = Lay::kStages, STEPS = Lay::kSteps;
  constexpr int TM = kWideTile, TN = kWideTile;
  constexpr int A_BYTES = TM * RB;
  constexpr int B_BYTES = TN * RB;
  constexpr int STAGE_BYTES = A_BYTES + B_BYTES;  // 32 KB or 64 KB
  constexpr int A_INST = A_BYTES / 1024 / 8;      // per wave: 2 or 4
  constexpr int B_INST = B_BYTES / 1024 / 8;
  constexpr int INST = A_INST + B_INST;
  __shared__ __align__(16) uint8_t st0[STAGE_BYTES];
  __shared__ __align__(16) uint8_t st1[STAGE_BYTES];
  __shared__ __align__(16) uint8_t st2[NS > 2 ? STAGE_BYTES : 16];
  __shared__ __align__(16) uint8_t st3[NS > 2 ? STAGE_BYTES : 16];
  __shared__ __align__(16) int32_t s_rnnz[TM];
  __shared__ float s_raux[TM];
  __shared__ float s_raux2[TM];
  __shared__ __align__(16) uint8_t s_rdeg[TM];
  __shared__ __align__(16) float s_rthr[SYM ? TM : 4];  // mirror screen thresholds
  __shared__ double s_rthrd[SYM ? TM : 1];                // mirror exact thresholds
  __shared__ int32_t s_mcnt[SYM ? TM : 1];                // mirror hits per reference
  __shared__ int32_t s_npool;                              // mirror hits past a column's slots
  // euclidean / pearson: the reference column's key weights as a query
  constexpr bool MW = SYM && MK != kKeyScore;
  __shared__ __align__(16) float s_rw1[MW ? TM : 4];
  __shared__ __align__(16) float s_rw2[MW ? TM : 4];
  // Jaccard: stored entries as a float, and the mirror screen's FMA
  // coefficients (jac_screen_coef) of the reference column as a query
  constexpr bool JC = sizeof(K) == 8;
  __shared__ __align__(16) float s_rnf[JC ? TM : 4];
  __shared__ __align__(16) float s_rja[JC && SYM ? TM : 4];
  __shared__ __align__(16) float s_rjb[JC && SYM ? TM : 4];

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = wave >> 2, wn = wave & 3;
  const int h = lane >> 5, n = lane & 31;
  int qt;
  int64_t rt;
  // the grid is two-dimensional when one dimension cannot hold every tile
  // (a dispatch counts work-items in 32 bits); dispatch order is x-major
  const int64_t blk = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
  if constexpr (SYM) {
    if (P.tri) {
      if (!map_tile_sym(blk, P.qtiles, P.rtiles, qt, rt)) return;
    } else {
      if (!map_tile(blk, (int64_t)gridDim.x * gridDim.y, P.qtiles, P.rtiles, qt, rt)) return;
    }
  } else {
    if (!map_tile(blk, (int64_t)gridDim.x * gridDim.y, P.qtiles, P.rtiles, qt, rt)) return;
  }
  const int64_t q_base = (int64_t)qt * TN;
  const int64_t r_base = rt * TM;
  const bool mirror = SYM && (!P.tri || rt > qt);  // workgroup-uniform
  // per-reference arrays of the epilogue, filled while the first chunks load
  auto fill_ref = [&]() {
  if (SYM && tid == 0) s_npool = 0;
  for (int t = tid; t < TM; t += NT) {
    const int64_t r = r_base + t;
    if (r < P.nref) {
      s_rnnz[t] = (int32_t)(P.ref_p[r + 1] - P.ref_p[r]);
      s_raux[t] = P.ref_aux ? P.ref_aux[r] : 0.f;
      s_raux2[t] = P.ref_aux2 ? P.ref_aux2[r] : 0.f;
      s_rdeg[t] = P.ref_deg ? P.ref_deg[r] : 0;
      if constexpr (SYM) {
        // reference r as a query (query position r): a float screen a little
        // below its threshold, as for the tile's own queries
        const double td = P.mthr[r];
        const float f = (float)td;
        s_rthr[t] = f - fabsf(f) * 9.5367431640625e-07f - 1e-30f;
        s_rthrd[t] = td;
        s_mcnt[t] = 0;
        if constexpr (MW) {
          const KeyWeights w =
              MK == kKeyPearson
                  ? pear_weights(s_raux[t], s_raux2[t], (uint32_t)s_rnnz[t], P.dot_e0, P.dot_e1,
                                 P.dot_nonneg)
                  : euc_weights((uint32_t)s_rnnz[t], P.dot_e0, P.dot_e1, P.dot_nonneg);
          s_rw1[t] = w.w1;
          s_rw2[t] = w.w2;
        }
        if constexpr (JC) {
          s_rja[t] = jac_screen_coef(s_rthr[t]);
          s_rjb[t] = s_rja[t] * (float)s_rnnz[t];
        }
      }
      if constexpr (JC) s_rnf[t] = (float)s_rnnz[t];
    } else {
      s_rnnz[t] = 0;
      s_raux[t] = 0.f;
      s_raux2[t] = 0.f;
      s_rdeg[t] = 1;
      if constexpr (SYM) {
        s_rthr[t] = INFINITY;
        s_rthrd[t] = INFINITY;
        s_mcnt[t] = 0;
      }
      if constexpr (MW) {
        s_rw1[t] = 0.f;
        s_rw2[t] = 0.f;
      }
      if constexpr (JC) {
        s_rnf[t] = 0.f;
        if constexpr (SYM) {
          s_rja[t] = __int_as_float(0x7fc00000);  // NaN: padding passes no screen
          s_rjb[t] = 0.f;
        }
      }
    }
  }
  };

  // Chunk c of a tile is one block of the chunk-major layout (rows 128 bytes
  // apart); a buffer resource per chunk and side covers the tile's rows of it
  // (rows past the end are clamped to the last row and feed only masked or
  // unused lanes).  Lane l of load i carries row kRowsPerInst i + l / kSlots
  // in physical slot l % kSlots, which holds logical segment
  // (l % kSlots) ^ swz(row).
  const int64_t a_rows = (P.nref - r_base < TM) ? P.nref - r_base : TM;
  const int64_t b_rows = (P.nqry - q_base < TN) ? P.nqry - q_base : TN;
  auto lane_offset = [&](int i, int64_t rows) {
    const int row = Lay::kRowsPerInst * i + lane / Lay::kSlots;
    const int seg = (lane % Lay::kSlots) ^ Lay::swz(row);
    const int r = row < rows ? row : (int)rows - 1;
    return (uint32_t)(r * kDenseChunk) + seg * 16;
  };
  uint32_t va[A_INST], vb[B_INST];
#pragma unroll
  for (int u = 0; u < A_INST; ++u) va[u] = lane_offset(wave * A_INST + u, a_rows);
#pragma unroll
  for (int u = 0; u < B_INST; ++u) vb[u] = lane_offset(wave * B_INST + u, b_rows);
  typedef __attribute__((address_space(3))) void lds_void;
  auto issue = [&](uint8_t* st, int64_t c) {
    const int kb = (int)(c * RB);
    const int kb_b = QS ? qseg_offset(kb, P.qseg) : kb;
    const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(P.ref_dense + dense_offset(r_base, kb, P.ref_ld)), 0, (int)(a_rows * kDenseChunk),
        kRsrcWord3);
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(P.qry_dense + dense_offset(q_base, kb_b, P.qry_ld)), 0,
        (int)(b_rows * kDenseChunk), kRsrcWord3);
#pragma unroll
    for (int u = 0; u < A_INST; ++u)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          ra, (lds_void*)(st + (wave * A_INST + u) * 1024), 16, (int)va[u], 0, 0, 0);
#pragma unroll
    for (int u = 0; u < B_INST; ++u)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rb, (lds_void*)(st + A_BYTES + (wave * B_INST + u) * 1024), 16, (int)vb[u], 0, 0, 0);
  };

  typename Op::Acc acc[4][2];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[b][j][r] = 0;

  // MFMA step t: lanes with h supply logical segment 2t + h of their row;
  // every A and B row read here has swz(row) == swz(n) (the row offsets are
  // multiples of 32).
  const int swz = Lay::swz(n);
  struct Frag {
    uint4 a[4], b[2];
  };
  auto load_frag = [&](Frag& f, const uint8_t* st, int t) {
    const int slot = ((2 * t + h) ^ swz) * 16;
#pragma unroll
    for (int b = 0; b < 4; ++b)
      f.a[b] = *reinterpret_cast<const uint4*>(st + (wm * 128 + b * 32 + n) * RB + slot);
#pragma unroll
    for (int j = 0; j < 2; ++j)
      f.b[j] = *reinterpret_cast<const uint4*>(st + A_BYTES + (wn * 64 + j * 32 + n) * RB + slot);
  };
  auto mma_frag = [&](const Frag& f) {
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int j = 0; j < 2; ++j) Op::mma(acc[b][j], f.a[b], f.b[j]);
  };
  Frag f0, f1;

  // A padded column is a multiple of 256 bytes, so the chunks come in whole
  // groups of NS.  Refills past the last chunk reload the last chunk into a
  // stage nobody reads again, so every step issues and waits alike and no
  // LDS read sits behind a branch.
  const int64_t nchunks = P.kpad_bytes / RB;
  const int64_t last = nchunks - 1;
  issue(st0, 0);
  issue(st1, last < 1 ? last : 1);
  if constexpr (NS > 2) {
    issue(st2, last < 2 ? last : 2);
    issue(st3, last < 3 ? last : 3);
  }
  // the per-reference loads go out after the first chunks', so the two
  // latencies overlap (a tile of short rows is a few chunks long)
  __builtin_amdgcn_sched_barrier(0);
  fill_ref();
  wait_vm<(NS - 1) * INST>();
  lds_barrier();  // the per-reference arrays are visible and chunk 0 has landed
  load_frag(f0, st0, 0);
  // Chunk c in stage CUR (its step-0 fragments already requested): the steps
  // alternate fragment sets, each requesting the next step's fragments before
  // running its MFMAs.  Before the last step: wait for this wave's loads of
  // chunk c + 1 (those of chunks c + 2 .. c + NS - 1 may still be in flight);
  // barrier, after which every wave has read CUR and chunk c + 1 is visible;
  // refill CUR with chunk c + NS and request chunk c + 1's step 0 while the
  // last step's MFMAs run.
#define SD_WIDE_STEPS(CUR)              \
  load_frag(f1, CUR, 1);                \
  __builtin_amdgcn_sched_barrier(0);    \
  mma_frag(f0);                         \
  __builtin_amdgcn_sched_barrier(0);    \
  if constexpr (STEPS == 4) {           \
    load_frag(f0, CUR, 2);              \
    __builtin_amdgcn_sched_barrier(0);  \
    mma_frag(f1);                       \
    __builtin_amdgcn_sched_barrier(0);  \
    load_frag(f1, CUR, 3);              \
    __builtin_amdgcn_sched_barrier(0);  \
    mma_frag(f0);                       \
    __builtin_amdgcn_sched_barrier(0);  \
  }
#define SD_WIDE_CHUNK(CUR, NXT, C)                  \
  {                                                 \
    SD_WIDE_STEPS(CUR)                              \
    wait_vm<(NS - 2) * INST>();                     \
    lds_barrier();                                  \
    issue(CUR, (C) + NS < last ? (C) + NS : last);  \
    load_frag(f0, NXT, 0);                          \
    __builtin_amdgcn_sched_barrier(0);              \
    mma_frag(f1);                                   \
    __builtin_amdgcn_sched_barrier(0);              \
  }
#define SD_WIDE_LAST(CUR)  \
  {                        \
    SD_WIDE_STEPS(CUR)     \
    mma_frag(f1);          \
  }
  int64_t c = 0;
  if constexpr (NS == 4) {
    for (; c + 4 < nchunks; c += 4) {
      SD_WIDE_CHUNK(st0, st1, c)
      SD_WIDE_CHUNK(st1, st2, c + 1)
      SD_WIDE_CHUNK(st2, st3, c + 2)
      SD_WIDE_CHUNK(st3, st0, c + 3)
    }
    // the last group of four (chunk c sits in st0)
    SD_WIDE_CHUNK(st0, st1, c)
    SD_WIDE_CHUNK(st1, st2, c + 1)
    SD_WIDE_CHUNK(st2, st3, c + 2)
    SD_WIDE_LAST(st3)
  } else {
    for (; c + 2 < nchunks; c += 2) {
      SD_WIDE_CHUNK(st0, st1, c)
      SD_WIDE_CHUNK(st1, st0, c + 1)
    }
    // the last pair (chunk c sits in st0)
    SD_WIDE_CHUNK(st0, st1, c)
    SD_WIDE_LAST(st1)
  }
#undef SD_WIDE_LAST
#undef SD_WIDE_CHUNK
#undef SD_WIDE_STEPS
  wait_vm<0>();  // no refill is left in flight when the workgroup ends

  // ---- epilogue: straight from the accumulators, no exchange.  acc[b][j][r]
  // is the value of reference 128 wm + 32b + 8(r/4) + 4h + (r%4) for query
  // 64 wn + 32j + n: each lane tests its 128 values against the thresholds of
  // its two queries with plain compares (most fail), runs the exact checks
  // only on those that pass, and appends them with one atomic per (lane,
  // query) after counting, so a wave never waits on a memory round trip per
  // candidate.
  // Cosine keys are the slab values themselves, so a screen against the
  // least float at or above each threshold is the exact test.
  constexpr bool kExactScreen = MK == kKeyScore && sizeof(K) != 8;
  // the lane's two queries: 64 wn + 32 j + n
  int64_t qp[2], qc[2];
  bool qv[2];
  uint32_t qn[2];
  float qa[2], qa2[2], thr_f[2];
  float ja[2], jb[2], qnf[2];  // Jaccard screen: c >= fma(ja, ni, jb), jb = ja nj
  KeyWeights kw[2];
  double thr_d[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    qp[j] = q_base + wn * 64 + j * 32 + n;
    qv[j] = qp[j] < P.nqry;
    qc[j] = qv[j] ? (P.qlist ? (int64_t)P.qlist[qp[j]] : qp[j]) : -1;
    qn[j] = qv[j] ? (uint32_t)(P.qry_p[qc[j] + 1] - P.qry_p[qc[j]]) : 0u;
    qa[j] = (qv[j] && P.qry_aux) ? P.qry_aux[qc[j]] : 0.f;
    qa2[j] = (qv[j] && P.qry_aux2) ? P.qry_aux2[qc[j]] : 0.f;
    kw[j] = MK == kKeyPearson ? pear_weights(qa[j], qa2[j], qn[j], P.dot_e0, P.dot_e1, P.dot_nonneg)
                              : euc_weights(qn[j], P.dot_e0, P.dot_e1, P.dot_nonneg);
    thr_d[j] = qv[j] ? P.thr[qp[j]] : INFINITY;
    const float t = (float)thr_d[j];
    if constexpr (kExactScreen) {
      // the least float at or above the threshold: for a float key x,
      // x >= thr_f exactly when (double) x >= thr_d
      thr_f[j] = ((double)t < thr_d[j]) ? next_float_up(t) : t;
    } else {
      // float screen a little below the threshold; the exact test follows
      thr_f[j] = t - fabsf(t) * 9.5367431640625e-07f - 1e-30f;
    }
    qnf[j] = (float)qn[j];
    ja[j] = jac_screen_coef(thr_f[j]);
    jb[j] = ja[j] * qnf[j];
  }
  const int64_t rg0 = r_base + wm * 128 + 4 * h;
  // exact test of one value: reference rg (tile row rloc, ni stored entries)
  // for query j; lanes of invalid queries have an infinite threshold
  auto exact = [&](const float x, int j, int rloc, int ni, int64_t rg, double& kd) -> bool {
    if (s_rdeg[rloc]) return false;
    const bool self = P.self_mode && rg == qc[j] + P.self_off;
    if (self && !P.include_self) return false;
    if constexpr (sizeof(K) == 8) {
      const uint32_t cc = (uint32_t)x;
      const uint32_t d = (uint32_t)ni + qn[j] - cc;
      kd = (d == 0) ? 1.0 : key_to_double(KeyJ{cc, d});
    } else {
      float key;
      if constexpr (MK == kKeyEuclid) {
        key = self ? FLT_MAX  // distance 0
                   : euc_key(kw[j], qa[j], s_raux[rloc], x);
      } else if constexpr (MK == kKeyPearson) {
        key = self ? FLT_MAX  // restored self correlation
                   : pear_key(kw[j], qa[j], qa2[j], s_raux[rloc], s_raux2[rloc], x);
      } else {
        key = self ? 2.0f : x;  // restored self similarity
      }
      kd = (double)key;
    }
    return kd >= thr_d[j];
  };
  // cheap screen (may keep too much, never too little); the self value is
  // marked separately below
  auto screen = [&](const float x, int j, float nif, float ra, float rc) -> bool {
    if constexpr (sizeof(K) == 8) {
      return x >= fmaf(ja[j], nif, jb[j]);  // jac_screen_coef
    } else if constexpr (MK == kKeyEuclid) {
      return euc_key(kw[j], qa[j], ra, x) >= thr_f[j];
    } else if constexpr (MK == kKeyPearson) {
      return pear_key(kw[j], qa[j], qa2[j], ra, rc, x) >= thr_f[j];
    } else {
      return x >= thr_f[j];
    }
  };
  // value r (0..15, per lane) of block b, query j: a select tree on r's four
  // bits (4 bit tests, 15 selects) rather than a compare and select per value,
  // each select waiting on its own compare.  The tree selects between
  // shuffled halves of the accumulator vector: selects between elements of
  // one vector are folded back into an element read at a variable index,
  // which is lowered to the compare chain.
  auto value_of = [&](int b, int j, int r) -> float {
    typedef float v8f __attribute__((ext_vector_type(8)));
    typedef float v4f __attribute__((ext_vector_type(4)));
    typedef float v2f __attribute__((ext_vector_type(2)));
    const v16f a = acc[b][j];
    const v8f l0 = (r & 1) ? __builtin_shufflevector(a, a, 1, 3, 5, 7, 9, 11, 13, 15)
                           : __builtin_shufflevector(a, a, 0, 2, 4, 6, 8, 10, 12, 14);
    const v4f l1 = (r & 2) ? __builtin_shufflevector(l0, l0, 1, 3, 5, 7)
                           : __builtin_shufflevector(l0, l0, 0, 2, 4, 6);
    const v2f l2 = (r & 4) ? __builtin_shufflevector(l1, l1, 1, 3)
                           : __builtin_shufflevector(l1, l1, 0, 2);
    return (r & 8) ? l2.y : l2.x;
  };
  // Pass 1, unrolled over the 128 values: plain compares only, one bit per
  // value (16 b + r of pm[j]).  The exact test and the append run in loops
  // over the set bits (pass 2), so the epilogue's code stays a few thousand
  // instructions: unrolled over all 128 values it outgrew the instruction
  // cache and was fetched again for every tile.
  uint64_t pm[2] = {0ull, 0ull};
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int q4 = 0; q4 < 4; ++q4) {
      // the four references 8 q4 + 4h + 0..3 of block b are consecutive
      const int rl = wm * 128 + b * 32 + 8 * q4 + 4 * h;
      float4 nn = make_float4(0.f, 0.f, 0.f, 0.f);
      float4 ra4 = make_float4(0.f, 0.f, 0.f, 0.f), rc4 = ra4;
      if constexpr (sizeof(K) == 8) nn = *reinterpret_cast<const float4*>(&s_rnf[rl]);
      if constexpr (MK != kKeyScore) ra4 = *reinterpret_cast<const float4*>(&s_raux[rl]);
      if constexpr (MK == kKeyPearson) rc4 = *reinterpret_cast<const float4*>(&s_raux2[rl]);
      const float nis[4] = {nn.x, nn.y, nn.z, nn.w};
      const float ras[4] = {ra4.x, ra4.y, ra4.z, ra4.w};
      const float rcs[4] = {rc4.x, rc4.y, rc4.z, rc4.w};
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        bool pass[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) pass[e] = screen(acc[b][j][4 * q4 + e], j, nis[e], ras[e], rcs[e]);
        // A compare per value, its lane masks ORed on the scalar unit; the
        // per-lane bits are set only for groups in which some lane of the wave
        // passed (about one group in four at the thresholds' density), so most
        // values cost one VALU compare instead of compare, select and or.
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(pass[0] | pass[1] | pass[2] | pass[3]) != 0,
                             0)) {
          uint32_t bits = 0u;
#pragma unroll
          for (int e = 0; e < 4; ++e) bits |= pass[e] ? 1u << e : 0u;
          pm[j] |= (uint64_t)bits << (16 * b + 4 * q4);
        }
      }
      // one group at a time: hoisting every group's LDS reads would hold
      // 16 x 4 aux values in registers next to the accumulators
      __builtin_amdgcn_sched_barrier(0);
    }
  // the query's own column, kept by the self rule whatever its slab value
  if (P.self_mode && P.include_self) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int64_t u = qv[j] ? qc[j] + P.self_off - r_base - wm * 128 : -1;
      if (u >= 0 && u < 128 && (((u & 31) >> 2) & 1) == h) {
        const int w = (int)(u & 31);
        pm[j] |= 1ull << (16 * (int)(u >> 5) + (((w >> 3) << 2) | (w & 3)));
      }
    }
  }
  // pass 2: exact test of the marked values, then one atomic per query to
  // reserve buffer slots, then the writes
  int cnt[2] = {0, 0};
  if constexpr (kExactScreen) {
    // the marks are the exact test already; what remains is to drop
    // degenerate and padding references (flag s_rdeg) and, without
    // include_self, the query's own column.  A degenerate column's slab value
    // is 0, so the flags matter only where a threshold is at most 0 or the
    // tile runs past the last reference (wave-uniform, rare).
    if (__builtin_amdgcn_ballot_w64(thr_f[0] <= 0.f || thr_f[1] <= 0.f) != 0 ||
        r_base + TM > P.nref) {
      uint64_t dm = 0ull;
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4) {
          const uint32_t d4 =
              *reinterpret_cast<const uint32_t*>(&s_rdeg[wm * 128 + b * 32 + 8 * q4 + 4 * h]);
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if ((d4 >> (8 * e)) & 0xFFu) dm |= 1ull << (16 * b + 4 * q4 + e);
        }
      pm[0] &= ~dm;
      pm[1] &= ~dm;
    }
    if (P.self_mode) {
      // the own column goes without include_self, and when it is degenerate
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int64_t u = qv[j] ? qc[j] + P.self_off - r_base - wm * 128 : -1;
        if (u >= 0 && u < 128 && (((u & 31) >> 2) & 1) == h) {
          const int w = (int)(u & 31);
          if (!P.include_self || s_rdeg[wm * 128 + (int)u])
            pm[j] &= ~(1ull << (16 * (int)(u >> 5) + (((w >> 3) << 2) | (w & 3))));
        }
      }
    }
    cnt[0] = __popcll(pm[0]);
    cnt[1] = __popcll(pm[1]);
  } else {
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        uint32_t m = (uint32_t)(pm[j] >> (16 * b)) & 0xFFFFu, keep = 0u;
        while (m) {
          const int r = __builtin_ctz(m);
          m &= m - 1u;
          const int rloc = wm * 128 + b * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
          const int64_t rg = rg0 + b * 32 + 8 * (r >> 2) + (r & 3);
          double kd;
          if (exact(value_of(b, j, r), j, rloc, s_rnnz[rloc], rg, kd)) keep |= 1u << r;
        }
        pm[j] = (pm[j] & ~(0xFFFFull << (16 * b))) | ((uint64_t)keep << (16 * b));
        cnt[j] += __popc(keep);
      }
  }
  if (cnt[0] + cnt[1] != 0) {
    int at[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) at[j] = cnt[j] ? atomicAdd(&P.cnt[qp[j]], cnt[j]) : 0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        uint32_t m = (uint32_t)(pm[j] >> (16 * b)) & 0xFFFFu;
        while (m) {
          const int r = __builtin_ctz(m);
          m &= m - 1u;
          const int rloc = wm * 128 + b * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
          const int64_t rg = rg0 + b * 32 + 8 * (r >> 2) + (r & 3);
          double kd;
          exact(value_of(b, j, r), j, rloc, s_rnnz[rloc], rg, kd);
          if (at[j] < P.cap) {
            P.fkey[qp[j] * P.cap + at[j]] = kd;
            P.fidx[qp[j] * P.cap + at[j]] = (int32_t)rg;
          }
          ++at[j];
        }
      }
  }
  // Mirror screen (a self-search's tiles off the diagonal, every tile of a
  // cross search): reference column rg as a query of the mirrored search,
  // against its own threshold, with this lane's query column as its
  // candidate.  No self pair lies off the diagonal, nor in a cross search.
  if constexpr (SYM) {
    if (mirror) {
      // candidates are valid, non-degenerate query columns only.  The flag
      // joins each compare below: clearing the marks afterwards under a
      // divergent condition was dropped by the compiler (the flag load fed an
      // empty branch), which let degenerate columns through.
      bool cand_ok[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int64_t cq = qv[j] ? qc[j] : 0;  // in range for every lane
        const int dg = P.mdeg ? (int)P.mdeg[cq] : 0;
        cand_ok[j] = qv[j] & (dg == 0);
      }
      // marks of blocks 2w and 2w + 1 in pw[j][w] (bit 16 (b & 1) + r)
      uint32_t pw[2][2] = {{0u, 0u}, {0u, 0u}};
      // cosine: each group's thresholds are read one group ahead, so their LDS
      // latency overlaps the previous group's compares (the barrier at each
      // group's end keeps the compiler from moving the reads itself).  The
      // other keys read several operands per group and gain nothing from it.
      float4 t4n = make_float4(0.f, 0.f, 0.f, 0.f);
      if constexpr (MK == kKeyScore && sizeof(K) != 8)
        t4n = *reinterpret_cast<const float4*>(&s_rthr[wm * 128 + 4 * h]);
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4) {
          const int rl = wm * 128 + b * 32 + 8 * q4 + 4 * h;
          float4 t4 = make_float4(0.f, 0.f, 0.f, 0.f);
          float4 w14 = t4, w24 = t4, ra4 = t4, rc4 = t4, ja4 = t4, jb4 = t4;
          if constexpr (sizeof(K) == 8) {
            ja4 = *reinterpret_cast<const float4*>(&s_rja[rl]);
            jb4 = *reinterpret_cast<const float4*>(&s_rjb[rl]);
          } else if constexpr (MK == kKeyScore) {
            t4 = t4n;
            if (b * 4 + q4 < 15) {
              const int g = b * 4 + q4 + 1;
              t4n = *reinterpret_cast<const float4*>(
                  &s_rthr[wm * 128 + (g >> 2) * 32 + 8 * (g & 3) + 4 * h]);
            }
          } else {
            t4 = *reinterpret_cast<const float4*>(&s_rthr[rl]);
          }
          if constexpr (MW) {
            w14 = *reinterpret_cast<const float4*>(&s_rw1[rl]);
            w24 = *reinterpret_cast<const float4*>(&s_rw2[rl]);
            ra4 = *reinterpret_cast<const float4*>(&s_raux[rl]);
          }
          if constexpr (MW && MK == kKeyPearson) rc4 = *reinterpret_cast<const float4*>(&s_raux2[rl]);
          const float ts[4] = {t4.x, t4.y, t4.z, t4.w};
          const float jas[4] = {ja4.x, ja4.y, ja4.z, ja4.w};
          const float jbs[4] = {jb4.x, jb4.y, jb4.z, jb4.w};
          const float w1s[4] = {w14.x, w14.y, w14.z, w14.w};
          const float w2s[4] = {w24.x, w24.y, w24.z, w24.w};
          const float ras[4] = {ra4.x, ra4.y, ra4.z, ra4.w};
          const float rcs[4] = {rc4.x, rc4.y, rc4.z, rc4.w};
          // each value's bit is selected on its own and the group's four are
          // ORed together, so the marks add one dependent step per group
          // rather than two per value
          uint32_t c[2][4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int r = 4 * q4 + e;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              const float x = acc[b][j][r];
              bool pass;
              if constexpr (sizeof(K) == 8) {
                pass = x >= fmaf(jas[e], qnf[j], jbs[e]);  // jac_screen_coef
              } else if constexpr (MK == kKeyEuclid) {
                pass = euc_key(KeyWeights{w1s[e], w2s[e], 0.f}, ras[e], qa[j], x) >= ts[e];
              } else if constexpr (MK == kKeyPearson) {
                pass = pear_key(KeyWeights{w1s[e], w2s[e], 0.f}, ras[e], rcs[e], qa[j], qa2[j],
                                x) >= ts[e];
              } else {
                pass = x >= ts[e];
              }
              c[j][e] = (pass & cand_ok[j]) ? 1u << (16 * (b & 1) + r) : 0u;
            }
          }
#pragma unroll
          for (int j = 0; j < 2; ++j) pw[j][b >> 1] |= (c[j][0] | c[j][1]) | (c[j][2] | c[j][3]);
          // the group's compares complete here: otherwise they sink below
          // every group's LDS reads, whose results then spill
          asm volatile("" : "+v"(pw[0][b >> 1]), "+v"(pw[1][b >> 1]));
          __builtin_amdgcn_sched_barrier(0);
        }
      uint64_t pmr[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) pmr[j] = (uint64_t)pw[j][0] | ((uint64_t)pw[j][1] << 32);
      // Gather the marked pairs per reference column in LDS (the stage
      // buffers are free once every wave's refills have landed): kMirrorSlots
      // per column in stage 0, further ones in a pool in stage 1.  Then one
      // thread per column computes the keys, runs the exact test against the
      // column's threshold and reserves its buffer slots with one atomic, and
      // the pool is tested entry by entry.  The lanes compute no key here, so
      // the accumulators stay in registers.  A pool that fills up pushes the
      // column's count past the buffer size, which sends that column to the
      // full per-query pass.
      constexpr int kMirrorSlots = 16;
      constexpr int kPool = STAGE_BYTES / 16;
      static_assert(TM * kMirrorSlots * 12 <= STAGE_BYTES, "mirror lists fit one stage");
      float* m_x = reinterpret_cast<float*>(st0);
      int32_t* m_idx = reinterpret_cast<int32_t*>(st0 + TM * kMirrorSlots * 4);
      uint32_t* m_nnz = reinterpret_cast<uint32_t*>(st0 + TM * kMirrorSlots * 8);
      int4* pool = reinterpret_cast<int4*>(st1);  // {value bits, candidate, its nnz, column}
      // key of slab value x for reference column rloc as the query and this
      // lane's query column j as its candidate (euclidean, pearson; Jaccard
      // and cosine store the value itself)
      auto mirror_fkey = [&](float x, int rloc, int j) -> float {
        if constexpr (MK == kKeyEuclid) {
          return euc_key(KeyWeights{s_rw1[rloc], s_rw2[rloc], 0.f}, s_raux[rloc], qa[j], x);
        } else if constexpr (MK == kKeyPearson) {
          return pear_key(KeyWeights{s_rw1[rloc], s_rw2[rloc], 0.f}, s_raux[rloc], s_raux2[rloc],
                          qa[j], qa2[j], x);
        } else {
          return x;
        }
      };
      // exact key of stored value x (the slab value for Jaccard, with cn the
      // candidate's stored entries; the float key otherwise) of tile column rloc
      auto mirror_key = [&](float x, int rloc, uint32_t cn) -> double {
        if constexpr (sizeof(K) == 8) {
          const uint32_t cc = (uint32_t)x;
          const uint32_t d = (uint32_t)s_rnnz[rloc] + cn - cc;
          return (d == 0) ? 1.0 : key_to_double(KeyJ{cc, d});
        } else {
          return (double)x;
        }
      };
      lds_barrier();  // no wave reads the stages any more; refills have landed (vmcnt 0)
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          uint32_t m = (uint32_t)(pmr[j] >> (16 * b)) & 0xFFFFu;
          while (m) {
            const int r = __builtin_ctz(m);
            m &= m - 1u;
            const int rloc = wm * 128 + b * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
            const float x = mirror_fkey(value_of(b, j, r), rloc, j);
            const int slot = atomicAdd(&s_mcnt[rloc], 1);
            if (slot < kMirrorSlots) {
              m_x[rloc * kMirrorSlots + slot] = x;
              m_idx[rloc * kMirrorSlots + slot] = (int32_t)qc[j];
              m_nnz[rloc * kMirrorSlots + slot] = qn[j];
            } else {
              const int ps = atomicAdd(&s_npool, 1);
              if (ps < kPool)
                pool[ps] = make_int4(__float_as_int(x), (int)qc[j], (int)qn[j], rloc);
              else if (r_base + rloc < P.nref)
                atomicAdd(&P.mcnt[r_base + rloc], P.mcap + 1);
            }
          }
        }
      lds_barrier();
      if (tid < TM) {
        const int64_t rg = r_base + tid;
        const int nm = min(s_mcnt[tid], kMirrorSlots);
        if (nm > 0 && rg < P.nref) {
          const double t = s_rthrd[tid];
          int keep = 0;
          for (int e = 0; e < nm; ++e)
            keep += mirror_key(m_x[tid * kMirrorSlots + e], tid, m_nnz[tid * kMirrorSlots + e]) >= t;
          if (keep > 0) {
            int at = atomicAdd(&P.mcnt[rg], keep);
            for (int e = 0; e < nm; ++e) {
              const double kd =
                  mirror_key(m_x[tid * kMirrorSlots + e], tid, m_nnz[tid * kMirrorSlots + e]);
              if (kd < t) continue;
              if (at < P.mcap) {
                P.mfkey[rg * P.mcap + at] = kd;
                P.mfidx[rg * P.mcap + at] = m_idx[tid * kMirrorSlots + e];
              }
              ++at;
            }
          }
        }
      }
      const int np = min(s_npool, kPool);
      for (int e = tid; e < np; e += NT) {
        const int4 v = pool[e];
        const int64_t rg = r_base + v.w;
        if (rg >= P.nref) continue;
        const double kd = mirror_key(__int_as_float(v.x), v.w, (uint32_t)v.z);
        if (kd >= s_rthrd[v.w]) {
          const int at = atomicAdd(&P.mcnt[rg], 1);
          if (at < P.mcap) {
            P.mfkey[rg * P.mcap + at] = kd;
            P.mfidx[rg * P.mcap + at] = v.y;
          }
        }
      }
    }
  }
}

// ---- merge per-split lists: top-L by (key desc, idx asc) --------------------
__global__ __launch_bounds__(256) void k_merge_lists(const double* __restrict__ key,
                                                     const int32_t* __restrict__ idx, int nlists,
                                                     int64_t nqry, int L,
                                                     double* __restrict__ out_key,
                                                     int32_t* __restrict__ out_idx, int64_t q0) {
  __shared__ BlockTopKStorage<512> st;
  const int64_t q = q0 + blockIdx.x;
  BlockTopK<512> tk(st, L, 1);
  tk.run((int64_t)nlists * L, [&](int64_t t, double& v, int64_t& id) {
    const int64_t s = t / L, e = t % L;
    const int64_t at = (s * nqry + q) * L + e;
    const int32_t i = idx[at];
    v = key[at];
    id = i;
    return i >= 0;
  });
  __syncthreads();
  const int cnt = st.cnt;
  for (int t = threadIdx.x; t < L; t += blockDim.x) {
    out_key[q * L + t] = (t < cnt) ? st.val[t] : 0.0;
    out_idx[q * L + t] = (t < cnt) ? (int32_t)st.idx[t] : -1;
  }
}

// ---- Jaccard results straight from exact lists ------------------------------------
__global__ void k_finish_exact(const double* __restrict__ key, const int32_t* __restrict__ idx,
                               int64_t nqry, int L, int k, int dist,
                               const uint8_t* __restrict__ qdeg, int32_t* __restrict__ out_idx,
                               double* __restrict__ out_val, int64_t out_ld) {
  const int64_t e0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e0 >= nqry * k) return;
  const int64_t q = e0 % nqry, t = e0 / nqry;
  const int64_t e = t * out_ld + q;  // output entry (q, t)
  const int32_t i = idx[q * L + t];
  if (i < 0 || (qdeg && qdeg[q])) {
    out_idx[e] = kNaInt;
    out_val[e] = __longlong_as_double((long long)kNaRealBits);
  } else {
    const double sim = key[q * L + t];
    out_idx[e] = i + 1;
    out_val[e] = dist ? (1.0 - sim) : sim;
  }
}


// Top-L of each query's appended candidates (key desc, idx asc); queries
// whose buffer overflowed are listed for a full list-mode pass instead.
__global__ __launch_bounds__(256) void k_select(const double* __restrict__ fkey,
                                                const int32_t* __restrict__ fidx,
                                                const int32_t* __restrict__ cnt, int cap,
                                                int64_t nqry, int L, double* __restrict__ out_key,
                                                int32_t* __restrict__ out_idx,
                                                int32_t* __restrict__ over_list,
                                                int32_t* __restrict__ over_count,
                                                int min_count,
                                                const int32_t* __restrict__ qlist, int64_t b0) {
  __shared__ BlockTopKStorage<512> st;
  // list position b; its query's buffer q (listed mode: qlist[b])
  const int64_t b = b0 + blockIdx.x;
  const int64_t q = qlist ? (int64_t)qlist[b] : b;
  const int n = cnt[q];
  // overflowed, or fewer candidates than the threshold guarantees complete
  // lists for: the query is searched again by a full list pass.  The list
  // records the query (or, for a query list, the list position b).
  if (n > cap || n < min_count) {
    if (threadIdx.x == 0) {
      const int at = atomicAdd(over_count, 1);
      if (over_list) over_list[at] = (int32_t)(qlist ? b : q);
    }
    for (int t = threadIdx.x; t < L; t += blockDim.x) out_idx[b * L + t] = -1;
    return;
  }
  BlockTopK<512> tk(st, L, 1);
  tk.run((int64_t)n, [&](int64_t t, double& v, int64_t& id) {
    v = fkey[q * cap + t];
    id = fidx[q * cap + t];
    return true;
  });
  __syncthreads();
  const int c = st.cnt;
  for (int t = threadIdx.x; t < L; t += blockDim.x) {
    out_key[b * L + t] = (t < c) ? st.val[t] : 0.0;
    out_idx[b * L + t] = (t < c) ? (int32_t)st.idx[t] : -1;
  }
}

// Threshold per query from a list: its L-th key, or -inf when not full.
__global__ void k_threshold(const double* __restrict__ key, const int32_t* __restrict__ idx,
                            int64_t nqry, int L, double* __restrict__ thr) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= nqry) return;
  thr[q] = (idx[q * L + L - 1] >= 0) ? key[q * L + L - 1] : -INFINITY;
}

// out[j] = src[pos[j]] (query ids of listed positions).
__global__ void k_gather_i32(const int32_t* __restrict__ src, const int32_t* __restrict__ pos,
                             int64_t n, int32_t* __restrict__ out) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x)
    out[j] = src[pos[j]];
}

// Copy list rows j into rows rows[j] of the main lists.
__global__ void k_scatter_lists(const double* __restrict__ key, const int32_t* __restrict__ idx,
                                const int32_t* __restrict__ rows, int64_t n, int L,
                                double* __restrict__ out_key, int32_t* __restrict__ out_idx) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n * L) return;
  const int64_t j = e / L, t = e % L;
  out_key[(int64_t)rows[j] * L + t] = key[e];
  out_idx[(int64_t)rows[j] * L + t] = idx[e];
}

}  // namespace



namespace {

__device__ inline uint64_t wave_shfl_u64(uint64_t v, int src) {
  const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  return (uint64_t)(uint32_t)__shfl((int)lo, src) |
         ((uint64_t)(uint32_t)__shfl((int)hi, src) << 32);
}

// Wave-level top-L (L <= 64) of the candidate buffers, one wave per list
// position, in BlockTopK's order: (order key, index) with the lower index
// winning ties, non-finite values never admitted.  Lane i holds the i-th best
// entry so far.  Candidates arrive 64 at a time; a chunk with none better
// than the current L-th is skipped, otherwise it is sorted across the wave
// (bitonic), merged with the list (the elementwise minimum against the
// reversed chunk is a bitonic sequence holding the 64 best of both) and
// cleaned back into order; a chunk with at most 12 such entries inserts them
// one by one at their ranks instead.  Lanes below L always hold the L best
// so far; the lanes above them stay in order.  Over-full and short buffers are listed for a full
// pass as in k_select.
__global__ __launch_bounds__(256) void k_select_wave(
    const double* __restrict__ fkey, const int32_t* __restrict__ fidx,
    const int32_t* __restrict__ cnt, int cap, int64_t nqry, int L, double* __restrict__ out_key,
    int32_t* __restrict__ out_idx, int32_t* __restrict__ over_list,
    int32_t* __restrict__ over_count, int min_count, const int32_t* __restrict__ qlist) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * 4 + wave;
  if (b >= nqry) return;  // whole waves leave; no barrier follows
  const int64_t q = qlist ? (int64_t)qlist[b] : b;
  const int n = cnt[q];
  if (n > cap || n < min_count) {
    if (lane == 0) {
      const int at = atomicAdd(over_count, 1);
      if (over_list) over_list[at] = (int32_t)(qlist ? b : q);  // as k_select
    }
    if (lane < L) out_idx[b * L + lane] = -1;
    return;
  }
  auto less = [](uint64_t ka, uint32_t ia, uint64_t kb, uint32_t ib) {
    return ka < kb || (ka == kb && ia < ib);
  };
  constexpr uint64_t kNone = ~0ULL;  // no finite value maps to it
  uint64_t lk = kNone;
  uint32_t li = 0xFFFFFFFFu;
  double lv = 0.0;
  const double* kb = fkey + q * cap;
  const int32_t* ib = fidx + q * cap;
  for (int base = 0; base < n; base += 64) {
    const int t = base + lane;
    uint64_t ck = kNone;
    uint32_t ci = 0xFFFFFFFFu;
    double cv = 0.0;
    if (t < n) {
      const double v = kb[t];
      if (isfinite(v)) {
        cv = v;
        ck = order_key(v, 1);
        ci = (uint32_t)ib[t];
      }
    }
    const uint64_t tk = wave_shfl_u64(lk, L - 1);
    const uint32_t ti = (uint32_t)__shfl((int)li, L - 1);
    if (!less(ck, ci, tk, ti)) {
      ck = kNone;
      ci = 0xFFFFFFFFu;
    }
    const uint64_t pass = __ballot(ck != kNone);
    if (pass == 0) continue;
    if (__popcll(pass) <= 12) {
      // a few entries (every chunk after the first two or so): insert each
      // at its rank in the list, the entries from that rank on moving up one
      // lane, which costs less than sorting and merging the chunk
      uint64_t m = pass;
      const int prev = lane > 0 ? lane - 1 : 0;
      while (m) {
        const int src = (int)__builtin_ctzll(m);
        m &= m - 1;
        const uint64_t nk = wave_shfl_u64(ck, src);
        const uint32_t ni = (uint32_t)__shfl((int)ci, src);
        const double nv = __shfl(cv, src);
        const int rank = (int)__popcll(__ballot(less(lk, li, nk, ni)));
        if (rank >= L) continue;  // no longer among the L best (wave-uniform)
        const uint64_t pk = wave_shfl_u64(lk, prev);
        const uint32_t pi = (uint32_t)__shfl((int)li, prev);
        const double pv = __shfl(lv, prev);
        if (lane > rank) {
          lk = pk;
          li = pi;
          lv = pv;
        } else if (lane == rank) {
          lk = nk;
          li = ni;
          lv = nv;
        }
      }
      continue;
    }
    // sort the chunk ascending across the wave
    for (int size = 2; size <= 64; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        const int partner = lane ^ stride;
        const uint64_t pk = wave_shfl_u64(ck, partner);
        const uint32_t pi = (uint32_t)__shfl((int)ci, partner);
        const double pv = __shfl(cv, partner);
        const bool lo = (lane & stride) == 0;
        const bool asc = (lane & size) == 0;
        const bool mine_less = less(ck, ci, pk, pi);
        if (mine_less != (lo == asc)) {
          ck = pk;
          ci = pi;
          cv = pv;
        }
      }
    }
    // the 64 best of list and chunk: elementwise minimum against the
    // reversed chunk, then a bitonic clean into ascending order
    {
      const uint64_t rk = wave_shfl_u64(ck, 63 - lane);
      const uint32_t ri = (uint32_t)__shfl((int)ci, 63 - lane);
      const double rv = __shfl(cv, 63 - lane);
      if (less(rk, ri, lk, li)) {
        lk = rk;
        li = ri;
        lv = rv;
      }
    }
    for (int stride = 32; stride > 0; stride >>= 1) {
      const int partner = lane ^ stride;
      const uint64_t pk = wave_shfl_u64(lk, partner);
      const uint32_t pi = (uint32_t)__shfl((int)li, partner);
      const double pv = __shfl(lv, partner);
      const bool lo = (lane & stride) == 0;
      const bool mine_less = less(lk, li, pk, pi);
      if (mine_less != lo) {
        lk = pk;
        li = pi;
        lv = pv;
      }
    }
  }
  if (lane < L) {
    const bool valid = lk != kNone;
    out_key[b * L + lane] = valid ? lv : 0.0;
    out_idx[b * L + lane] = valid ? (int32_t)li : -1;
  }
}

// One wave per query: exact FP64 values of the L candidates (the reference's
// cosine arithmetic), exact order, certificate against the best value any
// excluded column could have.
//
// Each wave first builds an LDS membership bitmap of its query's rows with
// per-word prefix counts (and stages the query's unit values), so a lane
// walks only its candidate column and looks each row up: the products of
// the common rows are added in ascending row order, exactly the order of the
// reference's merge walk (src/sparseDist.cpp:294-298).  When the bitmap does
// not fit (very tall matrices) the lane runs the merge walk itself.
// LDS per wave: the query's row bitmap (one bit per row) and the exclusive
// prefix of its per-word popcounts, which turn a row into the position of
// the query's value; sized to the row count so that many waves share a CU.
constexpr int kRerankWaveMax = 16384;
__host__ __device__ inline int rerank_wave_bytes(int nwords) {
  return ((((nwords + 3) & ~3) * 4 + ((nwords + 7) & ~7) * 2) + 15) & ~15;
}

// kRows16: the candidates' row indices are read from the uint16 copy (a
// separate instantiation, so the int32 walk keeps its register count and
// occupancy)
template <bool kRows16>
__global__ __launch_bounds__(256) void k_rerank(RerankArgs R) {
  extern __shared__ __align__(16) uint8_t rsm[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t j = blockIdx.x * 4 + wave;  // position in the query list
  const bool active = j < R.nqry;
  const int L = R.L;
  const int64_t jq = active ? (R.qlist ? (int64_t)R.qlist[j] : j) : 0;  // query column
  const int64_t q0 = R.ex.qry.p[jq], qn = active ? R.ex.qry.p[jq + 1] - q0 : 0;
  const int nwords = R.map_words;
  uint32_t* bm = reinterpret_cast<uint32_t*>(rsm + wave * rerank_wave_bytes(nwords));
  uint16_t* pre = reinterpret_cast<uint16_t*>(bm + ((nwords + 3) & ~3));
  const bool use_map = nwords > 0 && active && qn < 65536;
  const double* qv = R.ex.qry_unit + q0;  // read through the caches
  for (int w = lane; w < nwords; w += 64) bm[w] = 0u;
  __syncthreads();
  if (use_map) {
    for (int64_t e = lane; e < qn; e += 64) {
      const int32_t r = R.ex.qry.i[q0 + e];
      atomicOr(&bm[r >> 5], 1u << (r & 31));
    }
  }
  __syncthreads();
  if (nwords > 0) {
    // exclusive prefix of per-word popcounts: lane owns a contiguous run
    const int per = (nwords + 63) / 64;
    const int w0 = lane * per, w1 = (w0 + per < nwords) ? w0 + per : nwords;
    int local = 0;
    for (int w = w0; w < w1; ++w) local += __popc(bm[w]);
    int incl = local;
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(incl, off);
      if (lane >= off) incl += y;
    }
    int run = incl - local;
    for (int w = w0; w < w1; ++w) {
      pre[w] = (uint16_t)run;
      run += __popc(bm[w]);
    }
  }
  __syncthreads();
  if (!active) return;
  const int32_t ci = (lane < L) ? R.cand_idx[j * L + lane] : -1;
  const bool deg_q = R.qdeg && R.qdeg[jq];
  double c = -2.0, v = 0.0;
  const bool valid = ci >= 0 && !deg_q;
  if (valid) {
    if (R.ex.self_mode && R.ex.include_self && ci == jq + R.ex.self_off) {
      c = 1.0;  // restored self value (R/sparseKNN.R:270-279)
    } else {
      const int64_t a0 = R.ex.ref.p[ci], na = R.ex.ref.p[ci + 1] - a0;
      const double sa = R.ex.ref_scale[ci], sb = R.ex.qry_scale[jq];
      if (use_map) {
        const bool ea = (sa == 0.0), eb = (sb == 0.0);
        if (ea || eb) {
          c = (ea && eb) ? 1.0 : 0.0;
        } else {
          const int32_t* ra = R.ex.ref.i + a0;
          const double* ua = R.ex.ref_unit + a0;
          double acc = 0.0;
          if (na == R.ex.ref.nrow && qn == na) {
            // both columns store every row: the merge visits rows 0 .. n-1
            // in order, so the FP64 sum runs over the value arrays directly
            // (the same products in the same order; no index reads)
#pragma unroll 8
            for (int64_t e = 0; e < na; ++e) acc += ua[e] * qv[e];
          } else {
            // entries in ascending row order, so the FP64 sum keeps the
            // reference's order; the row indices stream in 16-byte loads (the
            // index array is 256-byte aligned, so entry a0 + e is 16-byte
            // aligned when (a0 + e) % 4 == 0) and a unit value is read only for
            // a row the query also stores
            auto step = [&](int32_t r, int64_t e) {
              const uint32_t word = bm[r >> 5];
              const uint32_t bit = 1u << (r & 31);
              if (word & bit) acc += ua[e] * qv[pre[r >> 5] + __popc(word & (bit - 1u))];
            };
            int64_t e = 0;
            if constexpr (kRows16) {
              // the same walk over the uint16 copy of the row indices (8 per
              // 16-byte load, entry a0 + e aligned when (a0 + e) % 8 == 0)
              const uint16_t* r16 = R.ex.ref_rows16 + a0;
              const int64_t head16 = min(na, (int64_t)((8 - (a0 & 7)) & 7));
              for (; e < head16; ++e) step(r16[e], e);
              for (; e + 64 <= na; e += 64) {
                const uint4* p64 = reinterpret_cast<const uint4*>(r16 + e);
                uint4 x[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) x[u] = p64[u];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                  const uint32_t w[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
#pragma unroll
                  for (int t = 0; t < 4; ++t) {
                    step((int32_t)(w[t] & 0xFFFFu), e + 8 * u + 2 * t);
                    step((int32_t)(w[t] >> 16), e + 8 * u + 2 * t + 1);
                  }
                }
              }
              for (; e + 8 <= na; e += 8) {
                const uint4 x = *reinterpret_cast<const uint4*>(r16 + e);
                const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                  step((int32_t)(w[t] & 0xFFFFu), e + 2 * t);
                  step((int32_t)(w[t] >> 16), e + 2 * t + 1);
                }
              }
              for (; e < na; ++e) step(r16[e], e);
            } else {
              const int64_t head = min(na, (int64_t)((4 - (a0 & 3)) & 3));
              for (; e < head; ++e) step(ra[e], e);
              // eight 16-byte loads in flight per lane before the walk uses them
              for (; e + 32 <= na; e += 32) {
                const int4* r32 = reinterpret_cast<const int4*>(ra + e);
                int4 x[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) x[u] = r32[u];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                  step(x[u].x, e + 4 * u);
                  step(x[u].y, e + 4 * u + 1);
                  step(x[u].z, e + 4 * u + 2);
                  step(x[u].w, e + 4 * u + 3);
                }
              }
              for (; e + 4 <= na; e += 4) {
                const int4 r4 = *reinterpret_cast<const int4*>(ra + e);
                step(r4.x, e);
                step(r4.y, e + 1);
                step(r4.z, e + 2);
                step(r4.w, e + 3);
              }
              for (; e < na; ++e) step(ra[e], e);
            }
          }
          c = acc;
          if (!(c == c)) c = 0.0;
          if (c > 1.0) c = 1.0;
          if (c < -1.0) c = -1.0;
        }
      } else {
        c = cos_value(R.ex.ref.i + a0, R.ex.ref_unit + a0, na, sa, R.ex.qry.i + q0,
                      R.ex.qry_unit + q0, qn, sb);
      }
    }
    v = R.ex.dist ? (1.0 - c) : c;
    if (v == 0.0) v = 0.0;
  }
  // sort lanes by (final value in the method's direction, idx)
  uint64_t key = valid ? order_key(v, R.decreasing) : ~0ULL;
  uint32_t id = valid ? (uint32_t)ci : 0xFFFFFFFFu;
  double cv = c, vv = v;
  for (int size = 2; size <= 64; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int partner = lane ^ stride;
      const uint64_t pk = wave_shfl_u64(key, partner);
      const uint32_t pi = (uint32_t)__shfl((int)id, partner);
      const double pc = __shfl(cv, partner);
      const double pv = __shfl(vv, partner);
      const bool is_lo = (lane & stride) == 0;
      const bool asc = (lane & size) == 0;
      const bool mine_less = key < pk || (key == pk && id < pi);
      const bool take_min = (is_lo == asc);
      const bool keep_mine = take_min ? mine_less : !mine_less;
      if (!keep_mine) {
        key = pk;
        id = pi;
        cv = pc;
        vv = pv;
      }
    }
  }
  // number of valid candidates and the list's worst approximate key
  const unsigned long long vmask = __ballot(valid);
  const int nvalid = __popcll(vmask);
  bool certified = true;
  if (nvalid >= L && R.k <= nvalid) {
    const double sL = R.cand_key[j * L + (L - 1)];
    const double ck = __shfl(cv, R.k - 1);
    const double ub = R.in_round;
    const double m = (double)qn;
    const double A = (2.0 * ub + ub * ub) + (1.0 + ub) * (1.0 + ub) * m * R.acc_round;
    double U;
    if (R.nonneg)
      U = (sL > 0.0) ? sL / (1.0 - A) : sL + A * 1e-300;
    else
      U = sL + A;
    U += 1e-12;
    certified = (A < 0.5) && (ck > U);
  }
  if (deg_q) certified = true;
  if (!certified) {
    if (lane == 0) {
      const int at = atomicAdd(R.fail_count, 1);
      R.fail_list[at] = (int32_t)jq;
    }
    return;
  }
  if (lane < R.k) {
    const int64_t e = (int64_t)lane * R.out_ld + jq;
    if (lane < nvalid && !deg_q) {
      R.out_idx[e] = (int32_t)id + 1;
      R.out_val[e] = vv;
    } else {
      R.out_idx[e] = kNaInt;
      R.out_val[e] = __longlong_as_double((long long)kNaRealBits);
    }
  }
}

// ---- euclidean / pearson re-rank ---------------------------------------------
// One wave per query, lane t < L takes candidate t and computes its exact
// FP64 value in the reference's arithmetic; lanes are sorted by (value in the
// method's direction, index).  Every reference outside the list has a key at
// most the list's last key, which bounds its true value:
//  * euclidean: the key is minus a lower bound LB of the squared distance D.
//    The reference's FP64 sum of squares is at least D (1 - rho) with
//    rho = (nnz(q) + nrow + 4) 2^-52 (at most nnz(q) + nrow non-negative
//    terms, each a rounded square of a rounded difference).  Certified when
//    the k-th squared distance num_k satisfies num_k (1 + 2^-49) < LB (1 - rho):
//    every outside value then exceeds num_k by more than a rounding of sqrt.
//  * pearson: the key is an upper bound UB of the correlation r.  The FP64
//    correlation differs from r by at most delta = (nnz(q) + 2 nrow + 32) 2^-50
//    (merge-walk and sum-of-squares rounding, relative to sq sy).  Certified
//    when the k-th correlation exceeds UB + delta + 1e-12 (the last term keeps
//    1 - r distinct after rounding in the distance form).
constexpr int kRerankEucQCap = 1024;  // query entries cached in LDS per wave

template <int M>
__global__ __launch_bounds__(256) void k_rerank_exact(RerankArgs R) {
  __shared__ int32_t s_qr[4][kRerankEucQCap];
  __shared__ double s_qv[4][kRerankEucQCap];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t j = blockIdx.x * 4 + wave;
  if (j >= R.nqry) return;  // whole waves leave; no barrier follows
  const int L = R.L;
  const int64_t jq = R.qlist ? (int64_t)R.qlist[j] : j;
  const int64_t q0 = R.ex.qry.p[jq], qn = R.ex.qry.p[jq + 1] - q0;
  const int32_t* qr = R.ex.qry.i + q0;
  const double* qv = R.ex.qry.x + q0;
  if (qn <= kRerankEucQCap) {
    for (int64_t e = lane; e < qn; e += 64) {
      s_qr[wave][e] = qr[e];
      s_qv[wave][e] = qv[e];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    qr = s_qr[wave];
    qv = s_qv[wave];
  }
  const int32_t ci = (lane < L) ? R.cand_idx[j * L + lane] : -1;
  const bool deg_q = R.qdeg && R.qdeg[jq];
  const bool valid = ci >= 0 && !deg_q;
  double c = 0.0, v = 0.0;  // euclidean: c = sum of squares; pearson: c = r
  if (valid) {
    if (M == SD_EUCLIDEAN) {
      const int64_t a0 = R.ex.ref.p[ci], na = R.ex.ref.p[ci + 1] - a0;
      c = euc_num(R.ex.ref.i + a0, R.ex.ref.x + a0, na, qr, qv, qn);
      v = sqrt(c);
    } else {
      if (R.ex.self_mode && R.ex.include_self && ci == jq + R.ex.self_off)
        c = 1.0;  // restored self value (R/sparseKNN.R:270-279)
      else
        c = pear_value(R.ex, ci, qr, qv, qn, jq);
      v = R.ex.dist ? (1.0 - c) : c;
      if (v == 0.0) v = 0.0;
    }
  }
  uint64_t key = valid ? order_key(v, R.decreasing) : ~0ULL;
  uint32_t id = valid ? (uint32_t)ci : 0xFFFFFFFFu;
  double cv = c, vv = v;
  for (int size = 2; size <= 64; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int partner = lane ^ stride;
      const uint64_t pk = wave_shfl_u64(key, partner);
      const uint32_t pi = (uint32_t)__shfl((int)id, partner);
      const double pc = __shfl(cv, partner);
      const double pv = __shfl(vv, partner);
      const bool is_lo = (lane & stride) == 0;
      const bool asc = (lane & size) == 0;
      const bool mine_less = key < pk || (key == pk && id < pi);
      const bool take_min = (is_lo == asc);
      const bool keep_mine = take_min ? mine_less : !mine_less;
      if (!keep_mine) {
        key = pk;
        id = pi;
        cv = pc;
        vv = pv;
      }
    }
  }
  const int nvalid = __popcll(__ballot(valid));
  bool certified = true;
  if (nvalid >= L && R.k <= nvalid) {
    const double lastkey = R.cand_key[j * L + (L - 1)];
    const double ck = __shfl(cv, R.k - 1);
    const double nrow = (double)R.ex.ref.nrow;
    if (M == SD_EUCLIDEAN) {
      const double lb = -lastkey;
      const double rho = ((double)qn + nrow + 4.0) * 0x1p-52;
      certified = lb > 0.0 && ck * (1.0 + 0x1p-49) < lb * (1.0 - rho);
    } else {
      const double delta = ((double)qn + 2.0 * nrow + 32.0) * 0x1p-50;
      certified = ck > lastkey + delta + 1e-12;
    }
  }
  if (deg_q) certified = true;
  if (!certified) {
    if (lane == 0) {
      const int at = atomicAdd(R.fail_count, 1);
      R.fail_list[at] = (int32_t)jq;
    }
    return;
  }
  if (lane < R.k) {
    const int64_t e = (int64_t)lane * R.out_ld + jq;
    if (lane < nvalid && !deg_q) {
      R.out_idx[e] = (int32_t)id + 1;
      R.out_val[e] = vv;
    } else {
      R.out_idx[e] = kNaInt;
      R.out_val[e] = __longlong_as_double((long long)kNaRealBits);
    }
  }
}

}  // namespace

// ----------------------------------------------------------------------------
// host launchers
// ----------------------------------------------------------------------------

template <class Op, class K, int L, int WM, int MK, bool QS>
static int launch_fused_t(sd_context* ctx, FusedKernelArgs& P) {
  const int64_t nblk = (int64_t)P.qtiles * P.splits;
  if (nblk * WM * 256 >= ((int64_t)1 << 32))
    return set_error(SD_EUNSUPPORTED, "slab grid too large (%lld workgroups)", (long long)nblk);
  hipLaunchKernelGGL((k_fused<Op, K, L, WM, MK, QS>), dim3((unsigned)nblk), dim3(WM * 256), 0,
                     ctx->stream, P);
  SD_HIP(hipGetLastError());
  return SD_OK;
}

// K bytes per row per chunk of the filtered pass (64 or 128)
constexpr int kFilterRowBytes = 128;

template <class Op, class K, int RB, int MK = kKeyScore, bool QS = false, bool SYM = false>
static int launch_filter_t(sd_context* ctx, FusedKernelArgs& P) {
  if (P.kpad_bytes <= 0 || P.kpad_bytes % (WideLayout<RB>::kStages * RB) != 0)
    return set_error(SD_EINVAL, "internal: padded row bytes %lld are not whole chunk groups",
                     (long long)P.kpad_bytes);
  P.qtiles = (int)((P.nqry + kWideTile - 1) / kWideTile);
  P.rtiles = (int)((P.nref + kWideTile - 1) / kWideTile);
  P.qgroups = (P.qtiles + 3) / 4;
  const bool tri = SYM && P.tri;
  if (tri && (P.nqry != P.nref || P.qtiles != P.rtiles))
    return set_error(SD_EINVAL, "internal: symmetric pass needs the query set to be the references");
  // slots: groups of 4 query tiles x their reference tiles (from the group's
  // first tile on, for the symmetric pass)
  const int64_t G = P.qgroups, R = P.rtiles;
  const int64_t need = tri ? 4 * G * R - 8 * G * (G - 1) : G * 4 * R;
  // x: a multiple of 8 (whole rounds of the XCD round-robin; 8 kSymChunk for
  // the symmetric order) small enough that x * 512 work-items fit a
  // dispatch's 32-bit grid size; y: the rest
  const int64_t unit = tri ? 8 * kSymChunk : 8;
  const int64_t gx = std::min<int64_t>((need + unit - 1) / unit * unit, (int64_t)1 << 20);
  const int64_t gy = (need + gx - 1) / gx;
  if (gy > 65535) return set_error(SD_EUNSUPPORTED, "slab grid too large (%lld tiles)", (long long)need);
  hipLaunchKernelGGL((k_filter_wide<Op, K, RB, MK, QS, SYM>), dim3((unsigned)gx, (unsigned)gy), dim3(512), 0,
                     ctx->stream, P);
  SD_HIP(hipGetLastError());
  return SD_OK;
}

template <class Op, class K, int WM, int MK = kKeyScore, bool QS = false>
static int launch_fused_L(sd_context* ctx, FusedKernelArgs& P, int L) {
  switch (L) {
    case 8: return launch_fused_t<Op, K, 8, WM, MK, QS>(ctx, P);
    case 16: return launch_fused_t<Op, K, 16, WM, MK, QS>(ctx, P);
    case 24: return launch_fused_t<Op, K, 24, WM, MK, QS>(ctx, P);
    case 32: return launch_fused_t<Op, K, 32, WM, MK, QS>(ctx, P);
    case 40: return launch_fused_t<Op, K, 40, WM, MK, QS>(ctx, P);
    case 48: return launch_fused_t<Op, K, 48, WM, MK, QS>(ctx, P);
    case 64: return launch_fused_t<Op, K, 64, WM, MK, QS>(ctx, P);
    default: return set_error(SD_EINVAL, "internal: unsupported list length %d", L);
  }
}

// Copy the densified columns of listed queries into a compact chunk-major
// block of n columns, position b holding column qlist[b] of the source (ld
// columns per chunk block); 8 lanes move one 128-byte chunk.
__global__ void k_gather_rows(const uint8_t* __restrict__ src, int64_t src_ld,
                              const int32_t* __restrict__ qlist, int64_t n, int64_t nchunks,
                              uint8_t* __restrict__ dst) {
  const int64_t total = n * nchunks * 8;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < total;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t part = v & 7, cell = v >> 3;
    const int64_t b = cell % n, c = cell / n;
    const int64_t kb = c * kDenseChunk + part * 16;
    *reinterpret_cast<uint4*>(dst + dense_offset(b, kb, n)) =
        *reinterpret_cast<const uint4*>(src + dense_offset(qlist[b], kb, src_ld));
  }
}

// Gathers a.qlist's query columns when a list is given; returns the layout
// to read and its columns per chunk block.
static int query_rows(sd_context* ctx, const FusedArgs& a, int64_t kpad_bytes,
                      const uint8_t** rows, int64_t* ld) {
  *rows = static_cast<const uint8_t*>(a.qry_dense);
  *ld = a.qry_ld;
  if (!a.qlist || a.nqry <= 0) return SD_OK;
  DevBuf& g = ctx->scratch[16];
  SD_TRY(g.ensure((size_t)a.nqry * (size_t)kpad_bytes + 64));
  const int64_t nchunks = kpad_bytes / kDenseChunk;
  hipLaunchKernelGGL(k_gather_rows, dim3(stride_grid(a.nqry * nchunks * 8)), dim3(256), 0,
                     ctx->stream, static_cast<const uint8_t*>(a.qry_dense), a.qry_ld, a.qlist,
                     a.nqry, nchunks, g.as<uint8_t>());
  SD_HIP(hipGetLastError());
  *rows = g.as<uint8_t>();
  *ld = a.nqry;
  return SD_OK;
}

bool fused_fits(int64_t row_bytes) {
  // one tile of 256 columns must stay below 2^31 bytes of buffer offset
  return row_bytes > 0 && row_bytes * 256 < ((int64_t)1 << 31);
}

int fused_list_len(int want) {
  static const int sizes[] = {8, 16, 24, 32, 40, 48, 64};
  for (int s : sizes)
    if (want <= s) return s;
  return 0;
}

int fused_slab_topl(sd_context* ctx, const FusedArgs& a, double* cand_key, int32_t* cand_idx) {
  FusedKernelArgs P{};
  P.ref_dense = static_cast<const uint8_t*>(a.ref_dense);
  P.qlist = a.qlist;
  P.nref = a.nref;
  P.nqry = a.nqry;
  P.kpad_bytes = dense_row_bytes(a.kind, a.kpad);
  if (!fused_fits(P.kpad_bytes))
    return set_error(SD_EINVAL, "internal: %lld padded rows exceed the slab kernel's tile range",
                     (long long)a.kpad);
  SD_TRY(query_rows(ctx, a, P.kpad_bytes, &P.qry_dense, &P.qry_ld));
  P.ref_ld = a.ref_ld;
  P.ref_p = a.ref_p;
  P.qry_p = a.qry_p;
  P.ref_aux = a.ref_aux;
  P.qry_aux = a.qry_aux;
  P.ref_aux2 = a.ref_aux2;
  P.qry_aux2 = a.qry_aux2;
  P.qseg = (int)dense_qseg_bytes(a.kind, a.kpad);
  P.dot_e0 = a.dot_e0;
  P.dot_e1 = a.dot_e1;
  P.dot_nonneg = a.dot_nonneg;
  P.ref_deg = a.ref_deg;
  P.self_mode = a.self_mode;
  P.self_off = a.self_off;
  P.include_self = a.include_self;
  P.method = a.method;
  constexpr int WM = 1;
  const int64_t TM = (int64_t)kTileRefs * WM;
  P.tile_stride = a.tile_stride < 1 ? 1 : a.tile_stride;
  P.ntiles = ((a.nref + TM - 1) / TM + P.tile_stride - 1) / P.tile_stride;
  P.qtiles = (int)((a.nqry + 255) / 256);
  int splits = a.splits < 1 ? 1 : a.splits;
  if (splits > P.ntiles) splits = (int)P.ntiles;
  P.tiles_per_split = (int)((P.ntiles + splits - 1) / splits);
  P.splits = (int)((P.ntiles + P.tiles_per_split - 1) / P.tiles_per_split);
  P.cand_key = cand_key;
  P.cand_idx = cand_idx;
  KernelTimer timer(ctx, &ctx->stats.slab_ms, &ctx->stats.slab_launches);
  if (a.method == SD_BINARY) {
    if (a.kind == kDenseF32Pattern) return launch_fused_L<OpF32, KeyJ, WM>(ctx, P, a.L);
    return launch_fused_L<OpF4, KeyJ, WM>(ctx, P, a.L);
  }
  if (a.prec == SD_PREC_FP32) {
    if (a.method == SD_EUCLIDEAN) return launch_fused_L<OpF32, float, WM, kKeyEuclid>(ctx, P, a.L);
    if (a.method == SD_PEARSON) return launch_fused_L<OpF32, float, WM, kKeyPearson>(ctx, P, a.L);
    return launch_fused_L<OpF32, float, WM>(ctx, P, a.L);
  }
  if (a.method == SD_EUCLIDEAN)
    return P.qseg ? launch_fused_L<OpBf16, float, WM, kKeyEuclid, true>(ctx, P, a.L)
                  : launch_fused_L<OpBf16, float, WM, kKeyEuclid>(ctx, P, a.L);
  if (a.method == SD_PEARSON)
    return P.qseg ? launch_fused_L<OpBf16, float, WM, kKeyPearson, true>(ctx, P, a.L)
                  : launch_fused_L<OpBf16, float, WM, kKeyPearson>(ctx, P, a.L);
  return launch_fused_L<OpBf16, float, WM>(ctx, P, a.L);
}

int fused_effective_splits(const FusedArgs& a) {
  const int64_t stride = a.tile_stride < 1 ? 1 : a.tile_stride;
  const int64_t ntiles = ((a.nref + kTileRefs - 1) / kTileRefs + stride - 1) / stride;
  int splits = a.splits < 1 ? 1 : a.splits;
  if (splits > ntiles) splits = (int)ntiles;
  const int64_t tps = (ntiles + splits - 1) / splits;
  return (int)((ntiles + tps - 1) / tps);
}

int fused_filter_pass(sd_context* ctx, const FusedArgs& a, const double* thr, int cap,
                      int32_t* cnt, double* fkey, int32_t* fidx) {
  FusedKernelArgs P{};
  P.ref_dense = static_cast<const uint8_t*>(a.ref_dense);
  P.qlist = a.qlist;
  P.nref = a.nref;
  P.nqry = a.nqry;
  P.kpad_bytes = dense_row_bytes(a.kind, a.kpad);
  if (!fused_fits(P.kpad_bytes))
    return set_error(SD_EINVAL, "internal: %lld padded rows exceed the slab kernel's tile range",
                     (long long)a.kpad);
  SD_TRY(query_rows(ctx, a, P.kpad_bytes, &P.qry_dense, &P.qry_ld));
  P.ref_ld = a.ref_ld;
  P.ref_p = a.ref_p;
  P.qry_p = a.qry_p;
  P.ref_aux = a.ref_aux;
  P.qry_aux = a.qry_aux;
  P.ref_aux2 = a.ref_aux2;
  P.qry_aux2 = a.qry_aux2;
  P.qseg = (int)dense_qseg_bytes(a.kind, a.kpad);
  P.dot_e0 = a.dot_e0;
  P.dot_e1 = a.dot_e1;
  P.dot_nonneg = a.dot_nonneg;
  P.ref_deg = a.ref_deg;
  P.self_mode = a.self_mode;
  P.self_off = a.self_off;
  P.include_self = a.include_self;
  P.method = a.method;
  P.qtiles = (int)((a.nqry + 255) / 256);
  P.rtiles = (int)((a.nref + kTileRefs - 1) / kTileRefs);
  P.tile_stride = 1;
  P.thr = thr;
  P.cap = cap;
  P.cnt = cnt;
  P.fkey = fkey;
  P.fidx = fidx;
  SD_TRY(zero_i32(ctx, cnt, a.nqry));
  // A self-search over all columns at once scores each unordered pair once:
  // the mirrored value is the same sum of products (on the split layout the
  // query side reads its segments swapped, so hi.lo and lo.hi are added in
  // the other order and the value may differ in the last bit, within the same
  // error bound), and each key uses its own query column's weights.
  const bool tri = a.self_mode && !a.qlist && a.nqry == a.nref && a.qry_dense == a.ref_dense &&
                   a.qry_p == a.ref_p && a.qry_aux == a.ref_aux && a.qry_aux2 == a.ref_aux2;
  // A cross search can feed the reverse search's buffers from the same tiles;
  // euclidean and pearson keys take the reverse query's weights from the
  // same dot error bound, which both directions must share.
  const bool cross = !tri && a.mirror && !a.qlist &&
                     (a.method == SD_COSINE || a.method == SD_BINARY ||
                      (a.mirror->e0 == a.dot_e0 && a.mirror->e1 == a.dot_e1 &&
                       a.mirror->nonneg == a.dot_nonneg));
  const bool sym = tri || cross;
  if (tri) {
    P.tri = 1;
    P.mthr = thr;
    P.mcap = cap;
    P.mcnt = cnt;
    P.mfkey = fkey;
    P.mfidx = fidx;
    P.mdeg = a.ref_deg;
  } else if (cross) {
    P.tri = 0;
    P.mthr = a.mirror->thr;
    P.mcap = a.mirror->cap;
    P.mcnt = a.mirror->cnt;
    P.mfkey = a.mirror->fkey;
    P.mfidx = a.mirror->fidx;
    P.mdeg = a.mirror->deg;
    SD_TRY(zero_i32(ctx, P.mcnt, a.nref));
    a.mirror->filled = true;
  }
  KernelTimer timer(ctx, &ctx->stats.filter_ms, &ctx->stats.filter_launches);
  ctx->stats.filter_pairs += tri ? a.nqry * (a.nqry + 1) / 2 : a.nqry * a.nref;
  constexpr int RB = kFilterRowBytes;
  if (a.method == SD_BINARY) {
    if (a.kind == kDenseF32Pattern)
      return sym ? launch_filter_t<OpF32, KeyJ, RB, kKeyScore, false, true>(ctx, P)
                 : launch_filter_t<OpF32, KeyJ, RB>(ctx, P);
    return sym ? launch_filter_t<OpF4, KeyJ, RB, kKeyScore, false, true>(ctx, P)
               : launch_filter_t<OpF4, KeyJ, RB>(ctx, P);
  }
  if (a.prec == SD_PREC_FP32) {
    if (a.method == SD_EUCLIDEAN)
      return sym ? launch_filter_t<OpF32, float, RB, kKeyEuclid, false, true>(ctx, P)
                 : launch_filter_t<OpF32, float, RB, kKeyEuclid>(ctx, P);
    if (a.method == SD_PEARSON)
      return sym ? launch_filter_t<OpF32, float, RB, kKeyPearson, false, true>(ctx, P)
                 : launch_filter_t<OpF32, float, RB, kKeyPearson>(ctx, P);
    return sym ? launch_filter_t<OpF32, float, RB, kKeyScore, false, true>(ctx, P)
               : launch_filter_t<OpF32, float, RB>(ctx, P);
  }
  if (a.method == SD_EUCLIDEAN) {
    if (sym)
      return P.qseg ? launch_filter_t<OpBf16, float, RB, kKeyEuclid, true, true>(ctx, P)
                    : launch_filter_t<OpBf16, float, RB, kKeyEuclid, false, true>(ctx, P);
    return P.qseg ? launch_filter_t<OpBf16, float, RB, kKeyEuclid, true>(ctx, P)
                  : launch_filter_t<OpBf16, float, RB, kKeyEuclid>(ctx, P);
  }
  if (a.method == SD_PEARSON) {
    if (sym)
      return P.qseg ? launch_filter_t<OpBf16, float, RB, kKeyPearson, true, true>(ctx, P)
                    : launch_filter_t<OpBf16, float, RB, kKeyPearson, false, true>(ctx, P);
    return P.qseg ? launch_filter_t<OpBf16, float, RB, kKeyPearson, true>(ctx, P)
                  : launch_filter_t<OpBf16, float, RB, kKeyPearson>(ctx, P);
  }
  return sym ? launch_filter_t<OpBf16, float, RB, kKeyScore, false, true>(ctx, P)
             : launch_filter_t<OpBf16, float, RB>(ctx, P);
}

int select_lists(sd_context* ctx, const double* fkey, const int32_t* fidx, const int32_t* cnt,
                 int cap, int64_t nqry, int L, double* out_key, int32_t* out_idx,
                 int32_t* over_list, int32_t* over_count, int min_count, const int32_t* qlist) {
  if (nqry <= 0) return SD_OK;
  if (L <= 64) {
    hipLaunchKernelGGL(k_select_wave, dim3((unsigned)((nqry + 3) / 4)), dim3(256), 0, ctx->stream,
                       fkey, fidx, cnt, cap, nqry, L, out_key, out_idx, over_list, over_count,
                       min_count, qlist);
    SD_HIP(hipGetLastError());
    return SD_OK;
  }
  for (int64_t b0 = 0; b0 < nqry; b0 += kMaxItemsPerLaunch) {
    const int64_t n = std::min(nqry - b0, kMaxItemsPerLaunch);
    hipLaunchKernelGGL(k_select, dim3((unsigned)n), dim3(256), 0, ctx->stream, fkey, fidx, cnt,
                       cap, nqry, L, out_key, out_idx, over_list, over_count, min_count, qlist, b0);
    SD_HIP(hipGetLastError());
  }
  return SD_OK;
}

int list_thresholds(sd_context* ctx, const double* key, const int32_t* idx, int64_t nqry, int L,
                    double* thr) {
  if (nqry <= 0) return SD_OK;
  hipLaunchKernelGGL(k_threshold, dim3((unsigned)((nqry + 255) / 256)), dim3(256), 0, ctx->stream,
                     key, idx, nqry, L, thr);
  SD_HIP(hipGetLastError());
  return SD_OK;
}

int scatter_lists(sd_context* ctx, const double* key, const int32_t* idx, const int32_t* rows,
                  int64_t n, int L, double* out_key, int32_t* out_idx) {
  if (n <= 0) return SD_OK;
  hipLaunchKernelGGL(k_scatter_lists, dim3((unsigned)((n * L + 255) / 256)), dim3(256), 0,
                     ctx->stream, key, idx, rows, n, L, out_key, out_idx);
  SD_HIP(hipGetLastError());
  return SD_OK;
}

int gather_i32(sd_context* ctx, const int32_t* src, const int32_t* pos, int64_t n, int32_t* out) {
  if (n <= 0) return SD_OK;
  hipLaunchKernelGGL(k_gather_i32, dim3(stride_grid(n)), dim3(256), 0, ctx->stream, src, pos, n, out);
  SD_HIP(hipGetLastError());
  return SD_OK;
}

int merge_lists(sd_context* ctx, const double* key, const int32_t* idx, int nlists, int64_t nqry,
                int L, double* out_key, int32_t* out_idx) {
  if (nqry <= 0) return SD_OK;
  for (int64_t q0 = 0; q0 < nqry; q0 += kMaxItemsPerLaunch) {
    const int64_t n = std::min(nqry - q0, kMaxItemsPerLaunch);
    hipLaunchKernelGGL(k_merge_lists, dim3((unsigned)n), dim3(256), 0, ctx->stream, key, idx,
                       nlists, nqry, L, out_key, out_idx, q0);
    SD_HIP(hipGetLastError());
  }
  return SD_OK;
}

int finish_exact_lists(sd_context* ctx, const double* key, const int32_t* idx, int64_t nqry, int L,
                       int k, int dist, const uint8_t* qdeg, int32_t* out_idx, double* out_val,
                       int64_t out_ld) {
  const int64_t n = nqry * k;
  if (n <= 0) return SD_OK;
  hipLaunchKernelGGL(k_finish_exact, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     key, idx, nqry, L, k, dist, qdeg, out_idx, out_val, out_ld);
  SD_HIP(hipGetLastError());
  return SD_OK;
}

int rerank_map_words(int64_t nrow) {
  const int64_t nwords = (nrow + 31) / 32;
  return (nwords < (1 << 20) && rerank_wave_bytes((int)nwords) <= kRerankWaveMax) ? (int)nwords
                                                                                : 0;
}

int rerank_cosine(sd_context* ctx, const RerankArgs& R) {
  if (R.nqry <= 0) return SD_OK;
  KernelTimer timer(ctx, &ctx->stats.rerank_ms, &ctx->stats.rerank_launches);
  const unsigned lds = R.map_words > 0 ? 4 * rerank_wave_bytes(R.map_words) : 0;
  if (R.ex.ref_rows16)
    hipLaunchKernelGGL(k_rerank<true>, dim3((unsigned)((R.nqry + 3) / 4)), dim3(256), lds,
                       ctx->stream, R);
  else
    hipLaunchKernelGGL(k_rerank<false>, dim3((unsigned)((R.nqry + 3) / 4)), dim3(256), lds,
                       ctx->stream, R);
  SD_HIP(hipGetLastError());
  return SD_OK;
}

template <int M>
static int rerank_exact(sd_context* ctx, const RerankArgs& R) {
  if (R.nqry <= 0) return SD_OK;
  if (R.L > 64) return set_error(SD_EINVAL, "internal: re-rank list of %d > 64", R.L);
  KernelTimer timer(ctx, &ctx->stats.rerank_ms, &ctx->stats.rerank_launches);
  hipLaunchKernelGGL(k_rerank_exact<M>, dim3((unsigned)((R.nqry + 3) / 4)), dim3(256), 0,
                     ctx->stream, R);
  SD_HIP(hipGetLastError());
  return SD_OK;
}

int rerank_euclid(sd_context* ctx, const RerankArgs& R) {
  return rerank_exact<SD_EUCLIDEAN>(ctx, R);
}

int rerank_pearson(sd_context* ctx, const RerankArgs& R) {
  return rerank_exact<SD_PEARSON>(ctx, R);
}

}  // namespace sd
