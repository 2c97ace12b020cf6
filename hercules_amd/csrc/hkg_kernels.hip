// Query kernels of libhkg for gfx950 (MI355X).
//
// Compiled with -ffp-contract=off: the squared distance and the iSAX lower
// bound must round every operation separately, in the reference's order
// (ts.c:83-109, sax.c:310-479), so their results are bit-identical to the
// reference's FP32 values.
//
// Layout (HBM, leaf order): X[n][L] f32, SAX[n][16] u8, per-leaf tables.
// A candidate is a 64-bit key  (d2 bits << 32) | public id : FP32 squared
// distances are non-negative, so the bit pattern orders like the value and
// sorting keys orders by (d2, id).
#include "hkg_kernels.h"

#include <float.h>

#include "hkg_ed.h"

#include <hipcub/hipcub.hpp>

namespace hkg {

__constant__ float c_bp[256];  // card-256 breakpoints, [255] unused

static const float h_bp[256] = {
#include "sax_bp256.inc"
    0.0f};

void host_breakpoints(float *out255)
{
  for (int i = 0; i < 255; ++i) out255[i] = h_bp[i];
}

hipError_t upload_breakpoints_dev()
{
  return hipMemcpyToSymbol(HIP_SYMBOL(c_bp), h_bp, sizeof(h_bp));
}

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// minidist_paa_to_isax_raw_SIMD for full-cardinality 16-symbol words
__device__ __forceinline__ float lb_sax_word(const float *paa, uint4 w, const float *bp, float ratio)
{
  uint32_t words[4] = {w.x, w.y, w.z, w.w};
  float t[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    int v = (words[i >> 2] >> ((i & 3) * 8)) & 0xFF;
    float lo = v == 0 ? -2000000.0f : bp[(v - 1) & 255];
    float hi = v == 255 ? 2000000.0f : bp[v];
    float p = paa[i];
    float r;
    if (lo > p) {
      float d = lo - p;
      r = d * d;
    } else if (hi < p) {
      float d = hi - p;
      r = d * d;
    } else {
      r = 0.0f;
    }
    t[i] = r;
  }
  float w0 = t[0] + t[8], w1 = t[1] + t[9], w2 = t[2] + t[10], w3 = t[3] + t[11];
  float w4 = t[4] + t[12], w5 = t[5] + t[13], w6 = t[6] + t[14], w7 = t[7] + t[15];
  float s = ((w0 + w1) + (w2 + w3)) + ((w4 + w5) + (w6 + w7));
  return s * ratio;
}

__device__ __forceinline__ bool lb_keep(float lb, uint32_t bound_bits)
{
  if (bound_bits >= kInfBits) return true;
  float b = __uint_as_float(bound_bits);
  return lb <= b * (1.0f + kLbRel) + kLbAbs;
}


// ---------------------------------------------------------------------------
// block top-k of distinct distances (256 threads)
// ---------------------------------------------------------------------------
struct TopK {
  uint64_t keys[kTopKCap];  // [0,cnt) sorted list, [256, 256+nstage) staged
  int cnt;
  int nstage;
  uint32_t kth_bits;        // bits of the k-th distinct d2, or kInfBits
};

__device__ __forceinline__ void topk_reset(TopK &s)
{
  if (threadIdx.x == 0) {
    s.cnt = 0;
    s.nstage = 0;
    s.kth_bits = kInfBits;
  }
}

// append one candidate (any thread); capacity 256 per merge round
__device__ __forceinline__ void topk_stage(TopK &s, uint64_t key)
{
  int slot = atomicAdd(&s.nstage, 1);
  s.keys[256 + slot] = key;
}

// all threads; call after a __syncthreads that follows the staging
__device__ void topk_merge(TopK &s, uint32_t *scan, int k)
{
  const int tid = threadIdx.x;
  const int n = s.cnt, m = s.nstage;
  if (m == 0) return;  // uniform: every thread read the same values
  for (int i = tid; i < kTopKCap; i += kBlock) {
    if ((i >= n && i < 256) || i >= 256 + m) s.keys[i] = ~0ull;
  }
  __syncthreads();
  // bitonic sort of 512 keys, one compare-exchange per thread per stage
  for (int size = 2; size <= kTopKCap; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      int i = ((tid / stride) * 2 * stride) + (tid % stride);
      int j = i + stride;
      uint64_t a = s.keys[i], b = s.keys[j];
      bool up = (i & size) == 0;
      if ((a > b) == up) {
        s.keys[i] = b;
        s.keys[j] = a;
      }
      __syncthreads();
    }
  }
  // keep the first key of every distinct distance (smallest id)
  uint32_t f0, f1;
  {
    int i0 = 2 * tid, i1 = 2 * tid + 1;
    uint64_t k0 = s.keys[i0], k1 = s.keys[i1];
    uint64_t km = i0 > 0 ? s.keys[i0 - 1] : ~0ull;
    f0 = (k0 != ~0ull) && (i0 == 0 || (uint32_t)(km >> 32) != (uint32_t)(k0 >> 32));
    f1 = (k1 != ~0ull) && ((uint32_t)(k0 >> 32) != (uint32_t)(k1 >> 32));
    scan[tid] = f0 + f1;
  }
  __syncthreads();
  // inclusive scan of 256 pair-counts (Hillis-Steele, small)
  for (int off = 1; off < kBlock; off <<= 1) {
    uint32_t v = tid >= off ? scan[tid - off] : 0;
    __syncthreads();
    scan[tid] += v;
    __syncthreads();
  }
  uint32_t excl = tid > 0 ? scan[tid - 1] : 0;
  uint32_t total = scan[kBlock - 1];
  uint64_t k0 = s.keys[2 * tid], k1 = s.keys[2 * tid + 1];
  __syncthreads();
  // compact into the upper half (it is free now), then move down
  if (f0 && excl < (uint32_t)k) s.keys[256 + excl] = k0;
  if (f1 && excl + f0 < (uint32_t)k) s.keys[256 + excl + f0] = k1;
  __syncthreads();
  int nc = total < (uint32_t)k ? (int)total : k;
  if (tid < nc) s.keys[tid] = s.keys[256 + tid];
  __syncthreads();
  if (tid == 0) {
    s.cnt = nc;
    s.nstage = 0;
    s.kth_bits = nc == k ? (uint32_t)(s.keys[k - 1] >> 32) : kInfBits;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// cooperative exact distances for a list of rows held in LDS
// ---------------------------------------------------------------------------
// Each (row, chunk) pair of a group is one lane-task: lanes of a wave cover
// consecutive chunks of a row (coalesced 32-byte reads), write the chunk
// sums into a padded LDS tile, then one lane per row folds its chunk sums
// in order (the reference's sequential accumulation).  Rows are staged as
// candidates when d2 <= bound (d2 equal to the bound is kept for the
// smallest-id rule).
template <int TILE>
struct RefineSharedT {
  float csum[kBlock / 64][TILE];  // per wave
};
using RefineShared = RefineSharedT<kRefineTile>;

// nrefined: lane 0 of every wave adds the rows its wave refined.
template <int TILE>
__device__ void refine_rows(const float *__restrict__ X, uint32_t L, const float *qs,
                            const uint64_t *rows, int nrows, const uint32_t *__restrict__ ids,
                            uint32_t bound_bits, TopK &tk, RefineSharedT<TILE> &rs, uint32_t &nrefined)
{
  const int cpr = (int)(L >> 3);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  float *cs = rs.csum[wave];
  if (cpr + 1 > TILE) {
    // rows longer than the tile: one row per wave at a time, its chunk sums
    // TILE at a time, folded in order into a running sum by lane 0 (every
    // wave runs the same number of rounds, so the block barriers match)
    for (int base = 0; base < nrows; base += kBlock / 64) {
      const int r = base + wave;
      const bool have = r < nrows;
      const uint64_t row = have ? rows[r] : 0;
      float acc = 0.0f;
      for (int c0 = 0; c0 < cpr; c0 += TILE) {
        const int nc = cpr - c0 < TILE ? cpr - c0 : TILE;
        if (have) {
          for (int c = lane; c < nc; c += 64) {
            const float *xp = X + row * L + (uint64_t)(c0 + c) * 8;
            const float4 xa = ld_stream(xp);
            const float4 xb = ld_stream(xp + 4);
            const float4 *qp = reinterpret_cast<const float4 *>(qs + (c0 + c) * 8);
            cs[c] = ed_chunk(qp[0], qp[1], xa, xb);
          }
        }
        __syncthreads();
        if (have && lane == 0)
          for (int c = 0; c < nc; ++c) acc = acc + cs[c];
        __syncthreads();
      }
      if (have && lane == 0) {
        if (__float_as_uint(acc) <= bound_bits) {
          const uint32_t id = ids ? ids[row] : (uint32_t)row;
          topk_stage(tk, make_key(acc, id));
        }
        ++nrefined;
      }
    }
    return;
  }
  // rows per wave group: the padded chunk tile must fit TILE floats
  int g = TILE / (cpr + 1);
  if (g > 64) g = 64;
  const int per_block = g * (kBlock / 64);
  for (int base = 0; base < nrows; base += per_block) {
    int wbeg = base + wave * g;
    int cnt = nrows - wbeg;
    if (cnt > g) cnt = g;
    if (cnt < 0) cnt = 0;
    const int tasks = cnt * cpr;
    for (int s = lane; s < tasks; s += 64) {
      int r = s / cpr;
      int c = s - r * cpr;
      uint64_t row = rows[wbeg + r];
      const float *xp = X + row * L + (uint64_t)c * 8;
      float4 xa = ld_stream(xp);
      float4 xb = ld_stream(xp + 4);
      const float4 *qp = reinterpret_cast<const float4 *>(qs + c * 8);
      cs[r * (cpr + 1) + c] = ed_chunk(qp[0], qp[1], xa, xb);
    }
    __syncthreads();
    if (lane < cnt) {
      const float *p = cs + lane * (cpr + 1);
      float acc = 0.0f;
      for (int c = 0; c < cpr; ++c) acc = acc + p[c];
      uint64_t row = rows[wbeg + lane];
      if (__float_as_uint(acc) <= bound_bits) {
        uint32_t id = ids ? ids[row] : (uint32_t)row;
        topk_stage(tk, make_key(acc, id));
      }
    }
    nrefined += lane == 0 ? (uint32_t)cnt : 0u;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// PAA (sax.c:46-56), prefix sums in FP64 for the EAPCA bound, ||q||^2 (FP64)
__global__ void __launch_bounds__(64) k_prep(const float *__restrict__ Q, uint32_t L,
                                             float *__restrict__ paa, double *__restrict__ psum,
                                             double *__restrict__ qn2)
{
  const uint32_t q = blockIdx.x;
  const float *x = Q + (uint64_t)q * L;
  const int lane = threadIdx.x;
  const int per = (int)(L / kPaaSegments);
  if (lane < kPaaSegments) {
    float a = 0.0f;
    for (int i = 0; i < per; ++i) a = a + x[lane * per + i];
    paa[q * kPaaSegments + lane] = a / (float)per;
  }
  if (lane == 0) {
    double *p1 = psum + (uint64_t)q * 2 * (L + 1);
    double *p2 = p1 + (L + 1);
    double s1 = 0.0, s2 = 0.0;
    p1[0] = 0.0;
    p2[0] = 0.0;
    for (uint32_t i = 0; i < L; ++i) {
      double v = (double)x[i];
      s1 += v;
      s2 += v * v;
      p1[i + 1] = s1;
      p2[i + 1] = s2;
    }
    qn2[q] = s2;
  }
}

// LB_EAPCA (hercules_node.c:879-924) for every (query, leaf): the query's
// per-segment mean and sd come from FP64 prefix sums instead of the
// reference's FP32 sums; the bound is used only with the certified slack.
// One leaf per thread, kEapcaQueries queries per block (grid.y: the query
// groups): a leaf's segment ends and ranges are read once for the group's
// queries, whose prefix sums sit in LDS (global memory when they do not
// fit); per query the arithmetic is the per-segment formula as it stands.
constexpr uint32_t kEapcaQueries = 8;
__global__ void __launch_bounds__(kBlock) k_eapca(uint32_t nleaf, uint32_t L, uint32_t nq,
                                                  const uint32_t *__restrict__ seg_off,
                                                  const int16_t *__restrict__ seg_end,
                                                  const float4 *__restrict__ ind,
                                                  const double *__restrict__ psum,
                                                  float *__restrict__ lb)
{
  extern __shared__ double sp[];
  const uint32_t q0 = blockIdx.y * kEapcaQueries;
  const uint32_t nqb = min(kEapcaQueries, nq - q0);
  const uint32_t stride = 2 * (L + 1);
  const double *gp = psum + (uint64_t)q0 * stride;
  const bool use_lds = kEapcaQueries * stride * sizeof(double) <= kEapcaLdsBytes;
  if (use_lds) {
    for (uint32_t i = threadIdx.x; i < nqb * stride; i += kBlock) sp[i] = gp[i];
    __syncthreads();
  }
  const double *pb = use_lds ? sp : gp;
  uint32_t leaf = blockIdx.x * kBlock + threadIdx.x;
  if (leaf >= nleaf) return;
  uint32_t b = seg_off[leaf], e = seg_off[leaf + 1];
  double sum[kEapcaQueries];
#pragma unroll
  for (uint32_t j = 0; j < kEapcaQueries; ++j) sum[j] = 0.0;
  int start = 0;
  for (uint32_t s = b; s < e; ++s) {
    int end = seg_end[s];
    double n = (double)(end - start);
    float4 g = ind[s];  // {max mean, min mean, max sd, min sd}
#pragma unroll
    for (uint32_t j = 0; j < kEapcaQueries; ++j) {
      if (j >= nqb) break;
      const double *p1 = pb + j * stride;
      const double *p2 = p1 + (L + 1);
      double mu = (p1[end] - p1[start]) / n;
      double var = (p2[end] - p2[start]) / n - mu * mu;
      double sd = var > 0.0 ? sqrt(var) : 0.0;
      double t = 0.0;
      if ((sd - g.z) * (sd - g.w) > 0.0) {
        double u = fmin(fabs(sd - (double)g.z), fabs(sd - (double)g.w));
        t += u * u;
      }
      if ((mu - g.x) * (mu - g.y) > 0.0) {
        double u = fmin(fabs(mu - (double)g.x), fabs(mu - (double)g.y));
        t += u * u;
      }
      sum[j] += t * n;
    }
    start = end;
  }
#pragma unroll
  for (uint32_t j = 0; j < kEapcaQueries; ++j)
    if (j < nqb) lb[(uint64_t)(q0 + j) * nleaf + leaf] = (float)sum[j];
}

// per query: the kBestLeaves leaves with the smallest bounds (ties: smaller
// leaf index first); missing ones (fewer leaves) are marked 0xFFFFFFFF
__global__ void __launch_bounds__(kBlock) k_best_leaf(uint32_t nleaf, const float *__restrict__ lb,
                                                      uint32_t *__restrict__ best)
{
  __shared__ uint64_t red[kBlock];
  const uint32_t q = blockIdx.x;
  uint64_t prev = 0;
  bool have_prev = false;
  for (int m = 0; m < kBestLeaves; ++m) {
    uint64_t mk = ~0ull;
    for (uint32_t i = threadIdx.x; i < nleaf; i += kBlock) {
      uint64_t key = ((uint64_t)__float_as_uint(lb[(uint64_t)q * nleaf + i]) << 32) | i;
      if ((!have_prev || key > prev) && key < mk) mk = key;
    }
    red[threadIdx.x] = mk;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) {
        uint64_t o = red[threadIdx.x + off];
        if (o < red[threadIdx.x]) red[threadIdx.x] = o;
      }
      __syncthreads();
    }
    const uint64_t found = red[0];
    __syncthreads();
    if (threadIdx.x == 0) best[q * kBestLeaves + m] = found == ~0ull ? 0xFFFFFFFFu : (uint32_t)found;
    prev = found;
    have_prev = true;
    if (found == ~0ull) {
      for (int r = m + 1; r < kBestLeaves; ++r)
        if (threadIdx.x == 0) best[q * kBestLeaves + r] = 0xFFFFFFFFu;
      break;
    }
  }
}

__device__ __forceinline__ void load_query(const float *__restrict__ Q, uint32_t q, uint32_t L, float *qs)
{
  const float4 *src = reinterpret_cast<const float4 *>(Q + (uint64_t)q * L);
  float4 *dst = reinterpret_cast<float4 *>(qs);
  for (uint32_t i = threadIdx.x; i < L / 4; i += kBlock) dst[i] = src[i];
}

// flush a block-local list into the query's global candidate buffer
__device__ void flush_topk(TopK &tk, int k, uint32_t q, uint32_t cap, uint64_t *__restrict__ cand,
                           uint32_t *__restrict__ cand_cnt, uint32_t *__restrict__ bound)
{
  __shared__ uint32_t s_base, s_bound;
  if (threadIdx.x == 0) {
    if (tk.cnt == k) atomicMin(&bound[q], tk.kth_bits);
    s_bound = __atomic_load_n(&bound[q], __ATOMIC_RELAXED);
    // count what survives the current global bound (the list is sorted)
    int nk = 0;
    while (nk < tk.cnt && (uint32_t)(tk.keys[nk] >> 32) <= s_bound) ++nk;
    s_base = nk ? atomicAdd(&cand_cnt[q], (uint32_t)nk) : 0;
    tk.nstage = nk;  // reuse as the flushed count
  }
  __syncthreads();
  int nk = tk.nstage;
  for (int i = threadIdx.x; i < nk; i += kBlock) {
    uint32_t slot = s_base + i;
    if (slot < cap) cand[(uint64_t)q * cap + slot] = tk.keys[i];
  }
  __syncthreads();
}

// approximate phase: the best leaves of every query, each scanned in full
// by its own block (grid.y = rank of the leaf)
__global__ void __launch_bounds__(kBlock) k_scan_best(const float *__restrict__ X, uint32_t L,
                                                      const float *__restrict__ Q,
                                                      const uint32_t *__restrict__ kq,
                                                      const uint32_t *__restrict__ best,
                                                      const uint64_t *__restrict__ leaf_pos,
                                                      const uint32_t *__restrict__ leaf_size,
                                                      const uint32_t *__restrict__ ids, uint32_t cap,
                                                      uint64_t *__restrict__ cand,
                                                      uint32_t *__restrict__ cand_cnt,
                                                      uint32_t *__restrict__ bound,
                                                      unsigned long long *__restrict__ counters)
{
  extern __shared__ float qs[];  // L floats
  __shared__ TopK tk;
  __shared__ RefineShared rs;
  __shared__ uint64_t rows[kBlock];
  __shared__ uint32_t scan[kBlock];
  const uint32_t q = blockIdx.x;
  const int k = (int)kq[q];
  const uint32_t leaf = best[q * kBestLeaves + blockIdx.y];
  if (leaf == 0xFFFFFFFFu) return;  // uniform for the block
  load_query(Q, q, L, qs);
  topk_reset(tk);
  const uint64_t pos = leaf_pos[leaf];
  const uint32_t size = leaf_size[leaf];
  uint32_t nref = 0;
  __syncthreads();
  for (uint32_t t0 = 0; t0 < size; t0 += kBlock) {
    uint32_t nr = size - t0 < (uint32_t)kBlock ? size - t0 : (uint32_t)kBlock;
    if (threadIdx.x < nr) rows[threadIdx.x] = pos + t0 + threadIdx.x;
    __syncthreads();
    refine_rows(X, L, qs, rows, (int)nr, ids, tk.kth_bits, tk, rs, nref);
    __syncthreads();
    topk_merge(tk, scan, k);
  }
  flush_topk(tk, k, q, cap, cand, cand_cnt, bound);
  if (threadIdx.x == 0) {
    atomicAdd(&counters[kCtrApprox], (unsigned long long)size);
    atomicAdd(&counters[kCtrApproxLeaves], 1ull);
  }
}

// Approximate phase of the BF16 path: in each of the query's best leaves the
// rows are ranked by xn2 - 2 q.x~ (q in FP32, x~ the row's BF16 copy: half
// the bytes of the rows), and only the M best-ranked ones get exact
// distances (the reference's order, as the full scan).  The k-th smallest
// of those is a k-th distance bound like the full scan's -- exact distances
// of real rows -- at worst less tight when the ranking misorders rows near
// the k-th.  Dynamic LDS: L floats of the query, then one ranking key per
// row of the leaf, a region the exact refinement reuses once the rows are
// chosen (so three blocks fit a CU).
constexpr uint32_t kApproxMargin = 32;  // rows refined beyond k
template <bool PRUNE>
__global__ void __launch_bounds__(kBlock) k_scan_best_bf16(const float *__restrict__ X,
                                                           const uint16_t *__restrict__ Xb,
                                                           const float *__restrict__ xn2, uint32_t L,
                                                           const float *__restrict__ Q,
                                                           const uint32_t *__restrict__ kq,
                                                           const uint8_t *__restrict__ sax,
                                                           const float *__restrict__ paa, uint32_t first_leaf,
                                                           const uint32_t *__restrict__ best,
                                                           const uint64_t *__restrict__ leaf_pos,
                                                           const uint32_t *__restrict__ leaf_size,
                                                           const uint32_t *__restrict__ ids, uint32_t cap,
                                                           uint64_t *__restrict__ cand,
                                                           uint32_t *__restrict__ cand_cnt,
                                                           uint32_t *__restrict__ bound,
                                                           unsigned long long *__restrict__ counters)
{
  extern __shared__ float dyn[];
  float *qs = dyn;
  uint32_t *key = reinterpret_cast<uint32_t *>(dyn + L);
  RefineShared &rs = *reinterpret_cast<RefineShared *>(dyn + L);  // after the keys are done with
  __shared__ TopK tk;
  __shared__ uint64_t rows[kBlock];
  __shared__ uint32_t scan[kBlock];
  __shared__ uint32_t hist[256];
  __shared__ uint32_t s_prefix, s_rank, s_nsel, s_ties, s_kept;
  __shared__ float sbp[PRUNE ? 256 : 1];
  __shared__ float spaa[kPaaSegments];
  const uint32_t q = blockIdx.x;
  const int k = (int)kq[q];
  const uint32_t leaf = best[q * kBestLeaves + first_leaf + blockIdx.y];
  if (leaf == 0xFFFFFFFFu) return;  // uniform for the block
  load_query(Q, q, L, qs);
  topk_reset(tk);
  if (PRUNE) {
    for (int i = threadIdx.x; i < 256; i += kBlock) sbp[i] = c_bp[i];
    if (threadIdx.x < kPaaSegments) spaa[threadIdx.x] = paa[q * kPaaSegments + threadIdx.x];
  }
  if (threadIdx.x == 0) s_kept = 0;
  const uint64_t pos = leaf_pos[leaf];
  const uint32_t size = leaf_size[leaf];
  // PRUNE: the query's k-th distance bound after its best leaf (an earlier
  // launch); a row whose iSAX lower bound is beyond it cannot enter the
  // top-k, so it is not ranked (its key stays the largest)
  const uint32_t tau = PRUNE ? __atomic_load_n(&bound[q], __ATOMIC_RELAXED) : kInfBits;
  const float ratio = (float)(L / kPaaSegments);
  __syncthreads();
  // ranking keys: one row per thread, read 256 bytes (16 loads) at a time;
  // the float is mapped to an unsigned key of the same order
  const uint32_t nc = L / 8;  // 16-byte chunks per row
  uint32_t kept = 0;
  for (uint32_t i = threadIdx.x; i < size; i += kBlock) {
    bool keep = true;
    if (PRUNE) {
      const uint4 w = *reinterpret_cast<const uint4 *>(sax + (pos + i) * kPaaSegments);
      keep = lb_keep(lb_sax_word(spaa, w, sbp, ratio), tau);
    }
    uint32_t kk = 0xFFFFFFFFu;
    if (keep) {
      const uint4 *xr = reinterpret_cast<const uint4 *>(Xb + (pos + i) * L);
      float dot = 0.0f;
      for (uint32_t c0 = 0; c0 < nc; c0 += 16) {
        uint4 v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = c0 + u < nc ? xr[c0 + u] : make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          if (c0 + u >= nc) break;
          const float4 qa = *reinterpret_cast<const float4 *>(qs + 8 * (c0 + u));
          const float4 qb = *reinterpret_cast<const float4 *>(qs + 8 * (c0 + u) + 4);
          dot = fmaf(qa.x, __uint_as_float(v[u].x << 16), dot);
          dot = fmaf(qa.y, __uint_as_float(v[u].x & 0xFFFF0000u), dot);
          dot = fmaf(qa.z, __uint_as_float(v[u].y << 16), dot);
          dot = fmaf(qa.w, __uint_as_float(v[u].y & 0xFFFF0000u), dot);
          dot = fmaf(qb.x, __uint_as_float(v[u].z << 16), dot);
          dot = fmaf(qb.y, __uint_as_float(v[u].z & 0xFFFF0000u), dot);
          dot = fmaf(qb.z, __uint_as_float(v[u].w << 16), dot);
          dot = fmaf(qb.w, __uint_as_float(v[u].w & 0xFFFF0000u), dot);
        }
      }
      const uint32_t b = __float_as_uint(fmaf(-2.0f, dot, xn2[pos + i]));
      kk = b ^ ((b >> 31) ? 0xFFFFFFFFu : 0x80000000u);
      ++kept;
    }
    key[i] = kk;
  }
  if (PRUNE) {
    for (int off = 32; off > 0; off >>= 1) kept += __shfl_down(kept, off);
    if ((threadIdx.x & 63) == 0 && kept) atomicAdd(&s_kept, kept);
  }
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&counters[kCtrApproxRanked], (unsigned long long)(PRUNE ? s_kept : size));
  if (PRUNE && s_kept == 0) {
    // every row of the leaf is beyond the bound: nothing to refine
    if (threadIdx.x == 0) {
      atomicAdd(&counters[kCtrApprox], (unsigned long long)size);
      atomicAdd(&counters[kCtrApproxLeaves], 1ull);
    }
    return;
  }
  // the M-th smallest key: radix select, 8 bits per pass from the top
  const uint32_t M = min(PRUNE ? s_kept : size, (uint32_t)min(k + (int)kApproxMargin, kBlock));
  if (threadIdx.x == 0) {
    s_prefix = 0;
    s_rank = M;
    s_nsel = 0;
    s_ties = 0;
  }
  for (int shift = 24; shift >= 0; shift -= 8) {
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t prefix = s_prefix, hi = shift == 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
    for (uint32_t i = threadIdx.x; i < size; i += kBlock) {
      const uint32_t kk = key[i];
      if ((kk & hi) == prefix) atomicAdd(&hist[(kk >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      // the first digit d (< 255) whose cumulative count reaches the rank:
      // lane l holds digits 4l..4l+3, a scan over the lanes finds the lane
      const uint32_t l = threadIdx.x, r = s_rank;
      const uint32_t h0 = hist[4 * l], h1 = hist[4 * l + 1], h2 = hist[4 * l + 2], h3 = hist[4 * l + 3];
      const uint32_t tot = h0 + h1 + h2 + h3;
      uint32_t inc = tot;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(inc, off);
        if (l >= (uint32_t)off) inc += v;
      }
      const uint64_t m = __ballot(inc >= r);
      const uint32_t at = m ? (uint32_t)(__ffsll((long long)m) - 1) : 63u;
      if (l == at) {
        uint32_t cum = inc - tot, d;
        if (cum + h0 >= r) {
          d = 4 * l;
        } else if (cum += h0, cum + h1 >= r) {
          d = 4 * l + 1;
        } else if (cum += h1, cum + h2 >= r) {
          d = 4 * l + 2;
        } else {
          cum += h2;
          d = 4 * l + 3;
        }
        s_prefix = prefix | (d << shift);
        s_rank = r - cum;
      }
    }
    __syncthreads();
  }
  // the rows below the M-th key, and as many equal to it as complete M
  const uint32_t kth = s_prefix, ties = s_rank;
  for (uint32_t i = threadIdx.x; i < size; i += kBlock) {
    const uint32_t kk = key[i];
    if (kk < kth || (kk == kth && atomicAdd(&s_ties, 1u) < ties)) rows[atomicAdd(&s_nsel, 1u)] = pos + i;
  }
  __syncthreads();
  uint32_t nref = 0;
  refine_rows(X, L, qs, rows, (int)s_nsel, ids, tk.kth_bits, tk, rs, nref);
  __syncthreads();
  topk_merge(tk, scan, k);
  flush_topk(tk, k, q, cap, cand, cand_cnt, bound);
  if (threadIdx.x == 0) {
    atomicAdd(&counters[kCtrApprox], (unsigned long long)size);
    atomicAdd(&counters[kCtrApproxLeaves], 1ull);
  }
}

// dense work list of every (query, leaf) pair: key = EAPCA bound bits when
// the pair survives the first bound, else all ones; the value is the pair's
// index q * nleaf + leaf.  Sorting by key gives a best-first order over all
// queries of the batch.
__global__ void __launch_bounds__(kBlock) k_collect(uint32_t nq, uint32_t nleaf,
                                                    const float *__restrict__ lb,
                                                    const uint32_t *__restrict__ best,
                                                    const uint32_t *__restrict__ bound,
                                                    uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                    uint32_t *__restrict__ nwork)
{
  const uint32_t q = blockIdx.y;
  const uint32_t leaf = blockIdx.x * kBlock + threadIdx.x;
  bool keep = false;
  const uint32_t idx = q * nleaf + leaf;
  if (leaf < nleaf) {
    float v = lb[idx];
    bool is_best = false;
#pragma unroll
    for (int m = 0; m < kBestLeaves; ++m) is_best |= leaf == best[q * kBestLeaves + m];
    keep = !is_best && lb_keep(v, bound[q]);
    keys[idx] = keep ? __float_as_uint(v) : 0xFFFFFFFFu;
    vals[idx] = idx;
  }
  uint64_t mask = __ballot(keep);
  if ((threadIdx.x & 63) == 0 && mask) atomicAdd(nwork, (uint32_t)__popcll(mask));
}

// persistent best-first (query, leaf) work queue: iSAX filter + exact
// refinement.  An item whose leaf bound no longer beats the query's current
// k-th distance is skipped when dequeued, and abandoned between tiles.
__global__ void __launch_bounds__(kBlock) k_filter_refine(
    const float *__restrict__ X, const uint8_t *__restrict__ sax, uint32_t L,
    const float *__restrict__ Q, const float *__restrict__ paa, const uint32_t *__restrict__ kq,
    const uint64_t *__restrict__ leaf_pos, const uint32_t *__restrict__ leaf_size,
    const uint32_t *__restrict__ ids, uint32_t nleaf, const float *__restrict__ lb,
    const uint32_t *__restrict__ work, const uint32_t *__restrict__ nwork_p, uint32_t *__restrict__ work_next,
    uint32_t cap, uint64_t *__restrict__ cand, uint32_t *__restrict__ cand_cnt, uint32_t *__restrict__ bound,
    unsigned long long *__restrict__ counters)
{
  extern __shared__ float qs[];
  __shared__ TopK tk;
  __shared__ RefineSharedT<kRefineTileFilter> rs;  // smaller: more blocks per CU
  __shared__ uint64_t rows[kFilterRows];
  __shared__ uint32_t scan[kBlock];
  __shared__ float bp[256];
  __shared__ float spaa[kPaaSegments];
  __shared__ uint32_t s_item, s_nsurv, s_bound, s_stop;
  __shared__ uint32_t s_chunk[kWorkChunk];
  __shared__ int s_nchunk, s_ci;
  for (int i = threadIdx.x; i < 256; i += kBlock) bp[i] = c_bp[i];
  const uint32_t nwork = *nwork_p;
  const float ratio = (float)(L / kPaaSegments);
  uint32_t nref = 0;
  unsigned long long nleaves = 0, nchecked = 0;
  if (threadIdx.x == 0) {
    s_nchunk = 0;
    s_ci = 0;
  }
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) {
      // next live item: take chunks of the sorted list, skip pruned pairs
      uint32_t found = 0xFFFFFFFFu;
      for (;;) {
        if (s_ci >= s_nchunk) {
          uint32_t base = atomicAdd(work_next, (uint32_t)kWorkChunk);
          if (base >= nwork) break;
          int m = (int)min((uint32_t)kWorkChunk, nwork - base);
          for (int j = 0; j < m; ++j) s_chunk[j] = work[base + j];
          s_nchunk = m;
          s_ci = 0;
        }
        uint32_t v = s_chunk[s_ci++];
        uint32_t qq = v / nleaf;
        if (lb_keep(lb[v], __atomic_load_n(&bound[qq], __ATOMIC_RELAXED))) {
          found = v;
          break;
        }
      }
      s_item = found;
    }
    __syncthreads();
    const uint32_t item = s_item;
    if (item == 0xFFFFFFFFu) break;
    const uint32_t q = item / nleaf, leaf = item - q * nleaf;
    const float item_lb = lb[item];
    const int k = (int)kq[q];
    load_query(Q, q, L, qs);
    if (threadIdx.x < kPaaSegments) spaa[threadIdx.x] = paa[q * kPaaSegments + threadIdx.x];
    topk_reset(tk);
    const uint64_t pos = leaf_pos[leaf];
    const uint32_t size = leaf_size[leaf];
    ++nleaves;
    // kFilterRows rows per step: each lane's iSAX words are loaded together,
    // the survivors refined kBlock at a time
    for (uint32_t t0 = 0; t0 < size; t0 += kFilterRows) {
      if (threadIdx.x == 0) {
        uint32_t gb = __atomic_load_n(&bound[q], __ATOMIC_RELAXED);
        s_bound = gb < tk.kth_bits ? gb : tk.kth_bits;
        s_nsurv = 0;
        s_stop = lb_keep(item_lb, gb) ? 0u : 1u;
      }
      __syncthreads();
      if (s_stop) break;
      const uint32_t bnd = s_bound;
      uint4 word[kFilterRows / kBlock];
#pragma unroll
      for (int u = 0; u < kFilterRows / kBlock; ++u) {
        const uint32_t r = t0 + u * kBlock + threadIdx.x;
        word[u] = r < size ? *reinterpret_cast<const uint4 *>(sax + (pos + r) * kPaaSegments) : make_uint4(0, 0, 0, 0);
      }
      const int lane = threadIdx.x & 63;
#pragma unroll
      for (int u = 0; u < kFilterRows / kBlock; ++u) {
        const uint32_t r = t0 + u * kBlock + threadIdx.x;
        const bool keep = r < size && lb_keep(lb_sax_word(spaa, word[u], bp, ratio), bnd);
        const uint64_t mask = __ballot(keep);
        const int leader = __ffsll((long long)mask) - 1;
        uint32_t wbase = 0;
        if (mask) {
          if (lane == leader) wbase = atomicAdd(&s_nsurv, (uint32_t)__popcll(mask));
          wbase = __shfl(wbase, leader);
        }
        if (keep) rows[wbase + __popcll(mask & ((1ull << lane) - 1))] = pos + r;
      }
      nchecked += (threadIdx.x == 0) ? (unsigned long long)min((uint32_t)kFilterRows, size - t0) : 0ull;
      __syncthreads();
      const int ns = (int)s_nsurv;
      for (int g = 0; g < ns; g += kBlock) {
        refine_rows(X, L, qs, rows + g, ns - g < kBlock ? ns - g : kBlock, ids, bnd, tk, rs, nref);
        __syncthreads();
        topk_merge(tk, scan, k);
      }
    }
    flush_topk(tk, k, q, cap, cand, cand_cnt, bound);
  }
  if ((threadIdx.x & 63) == 0 && nref) atomicAdd(&counters[kCtrRefined], (unsigned long long)nref);
  if (threadIdx.x == 0) {
    if (nleaves) atomicAdd(&counters[kCtrLeaves], nleaves);
    if (nchecked) atomicAdd(&counters[kCtrSaxChecked], nchecked);
  }
}

// PSCAN: exact full scan (parallel_scan_utils.c:349-429 does one query per
// worker over the original-order file).  A block serves a group of QN
// queries (QN <= kScanQ) and sweeps the rows persistently, one row per
// lane: the lane streams its row in 128-byte steps (four 8-float chunks)
// and folds the chunk sums in the reference's order (ts.c:83-109: an
// 8-term tree per chunk, chunks added in sequence), abandoning the row once
// every query's partial sum exceeds its bound (partial sums only grow, so
// no row within the bound is lost).  Queries are read from LDS by
// broadcast.  Block x of every query group reads the same rows (grid.x a
// multiple of the XCD count), so one group brings them from HBM and the
// others find them in the caches.
template <int QN>
__global__ void __launch_bounds__(kBlock) k_scan(const float *__restrict__ X, uint64_t n, uint32_t L,
                                                 const float *__restrict__ Q, uint32_t nq,
                                                 const uint32_t *__restrict__ kq,
                                                 const uint32_t *__restrict__ ids, uint32_t cap,
                                                 uint64_t *__restrict__ cand,
                                                 uint32_t *__restrict__ cand_cnt,
                                                 uint32_t *__restrict__ bound)
{
  extern __shared__ float qs[];  // QN * L
  __shared__ TopK tk[QN];
  __shared__ uint32_t scan[kBlock];
  const uint32_t q0 = blockIdx.y * QN;
  const int nqg = (int)(nq - q0 < (uint32_t)QN ? nq - q0 : (uint32_t)QN);
  for (int j = 0; j < nqg; ++j) load_query(Q, q0 + j, L, qs + j * L);
  for (int j = 0; j < QN; ++j) topk_reset(tk[j]);
  __syncthreads();
  const int nch = (int)(L >> 3);  // 8-float chunks per row
  const float4 *q4 = reinterpret_cast<const float4 *>(qs);
  for (uint64_t base = (uint64_t)blockIdx.x * kBlock; base < n; base += (uint64_t)gridDim.x * kBlock) {
    uint32_t bnd[QN];
#pragma unroll
    for (int j = 0; j < QN; ++j) {
      const uint32_t gb = j < nqg ? __atomic_load_n(&bound[q0 + j], __ATOMIC_RELAXED) : 0u;
      bnd[j] = gb < tk[j].kth_bits ? gb : tk[j].kth_bits;
    }
    const uint64_t row = base + threadIdx.x;
    if (row < n) {
      const float4 *xp = reinterpret_cast<const float4 *>(X + row * L);
      float acc[QN];
#pragma unroll
      for (int j = 0; j < QN; ++j) acc[j] = 0.0f;
      bool alive = true;
      int c = 0;
      // four chunks (128 bytes) per step
      for (; c + 4 <= nch && alive; c += 4) {
        float4 xv[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) xv[i] = ld_stream(reinterpret_cast<const float *>(xp + 2 * c + i));
        alive = false;
#pragma unroll
        for (int j = 0; j < QN; ++j) {
          if (j < nqg) {
#pragma unroll
            for (int sub = 0; sub < 4; ++sub)
              acc[j] = acc[j] + ed_chunk(q4[(j * L >> 2) + 2 * (c + sub)], q4[(j * L >> 2) + 2 * (c + sub) + 1],
                                         xv[2 * sub], xv[2 * sub + 1]);
            alive |= __float_as_uint(acc[j]) <= bnd[j];
          }
        }
      }
      for (; c < nch && alive; ++c) {
        const float4 xa = ld_stream(reinterpret_cast<const float *>(xp + 2 * c));
        const float4 xb = ld_stream(reinterpret_cast<const float *>(xp + 2 * c + 1));
        alive = false;
#pragma unroll
        for (int j = 0; j < QN; ++j) {
          if (j < nqg) {
            acc[j] = acc[j] + ed_chunk(q4[(j * L >> 2) + 2 * c], q4[(j * L >> 2) + 2 * c + 1], xa, xb);
            alive |= __float_as_uint(acc[j]) <= bnd[j];
          }
        }
      }
      if (alive) {  // every chunk was added: acc is the exact distance
        const uint32_t id = ids ? ids[row] : (uint32_t)row;
#pragma unroll
        for (int j = 0; j < QN; ++j)
          if (j < nqg && __float_as_uint(acc[j]) <= bnd[j]) topk_stage(tk[j], make_key(acc[j], id));
      }
    }
    __syncthreads();
    for (int j = 0; j < nqg; ++j) {
      const int k = (int)kq[q0 + j];
      topk_merge(tk[j], scan, k);
      // a full list's k-th distinct distance bounds the query's k-th: share
      // it at once, as the reference's workers share one k-th distance
      // (parallel_scan_utils.c:381-383), so every block prunes against the
      // tightest bound found anywhere so far
      if (threadIdx.x == 0 && tk[j].cnt == k && tk[j].kth_bits < bnd[j]) atomicMin(&bound[q0 + j], tk[j].kth_bits);
    }
  }
  for (int j = 0; j < nqg; ++j) flush_topk(tk[j], (int)kq[q0 + j], q0 + j, cap, cand, cand_cnt, bound);
}

// final per-query selection over the candidate buffer
__global__ void __launch_bounds__(kBlock) k_select(uint32_t kmax, const uint32_t *__restrict__ kq,
                                                   uint32_t cap, const uint64_t *__restrict__ cand,
                                                   const uint32_t *__restrict__ cand_cnt,
                                                   const uint64_t *__restrict__ idmap64,
                                                   uint64_t *__restrict__ out_ids,
                                                   float *__restrict__ out_d2,
                                                   uint32_t *__restrict__ overflow)
{
  __shared__ TopK tk;
  __shared__ uint32_t scan[kBlock];
  const uint32_t q = blockIdx.x;
  const int k = (int)kq[q];
  uint32_t cnt = cand_cnt[q];
  if (cnt > cap) {
    if (threadIdx.x == 0) atomicAdd(overflow, 1u);
    cnt = cap;
  }
  topk_reset(tk);
  __syncthreads();
  for (uint32_t t0 = 0; t0 < cnt; t0 += kBlock) {
    uint32_t i = t0 + threadIdx.x;
    if (i < cnt) {
      uint64_t key = cand[(uint64_t)q * cap + i];
      if ((uint32_t)(key >> 32) <= tk.kth_bits) topk_stage(tk, key);
    }
    __syncthreads();
    topk_merge(tk, scan, k);
  }
  for (uint32_t r = threadIdx.x; r < kmax; r += kBlock) {
    uint64_t id = ~0ull;
    float d = FLT_MAX;
    if ((int)r < tk.cnt) {
      uint64_t key = tk.keys[r];
      d = __uint_as_float((uint32_t)(key >> 32));
      uint32_t id32 = (uint32_t)key;
      id = idmap64 ? idmap64[id32] : (uint64_t)id32;
    }
    out_ids[(uint64_t)q * kmax + r] = id;
    out_d2[(uint64_t)q * kmax + r] = d;
  }
}

// Last stage of the BF16 path: one block per query merges the approximate
// stage's candidates and the exact distances of the survivors of the lower
// bound into the final top-k, refining survivors in kRefinePasses passes of
// increasing bound, guided by a histogram of their bounds: first the
// ~kRefineFirst smallest bounds (which fixes a k-th distance close to the
// final one), then about 4x as many, and so on (the last pass: the rest),
// each pass taking only survivors whose bound is within the current k-th
// distance.  Every survivor whose bound is within the final k-th distance is
// refined.  The survivors are every series whose bound is within `bound` (an
// upper bound of the k-th distance counted over series), so the answer is
// exact when no survivor was dropped for lack of room and the final k-th
// distinct distance is within `bound`; otherwise the query is flagged
// (scnt = ~0) and the caller answers it again, seeded with this k-th
// distance.  Rows are refined by whole waves (refine_rows): on MI355X at
// 100M x 256, 2048 queries, k ~ U{1..100}, two passes take 3.4 ms, four
// 4.5 ms, and four with one row per thread and early abandon 5.7 ms.
constexpr int kSurvBins = 256;
constexpr uint32_t kRefineFirst = 2048;
constexpr int kRefinePasses = 2;

__global__ void __launch_bounds__(kBlock) k_bf16_refine(
    const float *__restrict__ X, uint32_t L, const float *__restrict__ Q, const uint32_t *__restrict__ kq,
    uint32_t kmax, const uint32_t *__restrict__ ids, uint32_t cap, const uint64_t *__restrict__ cand,
    const uint32_t *__restrict__ cand_cnt, uint32_t scap, const uint64_t *__restrict__ surv,
    uint32_t *__restrict__ scnt, const float *__restrict__ thr0, const uint32_t *__restrict__ bound,
    const uint64_t *__restrict__ idmap64, uint64_t *__restrict__ out_ids, float *__restrict__ out_d2,
    uint32_t *__restrict__ overflow, unsigned long long *__restrict__ counters)
{
  extern __shared__ float qs[];
  __shared__ TopK tk;
  __shared__ RefineShared rs;
  __shared__ uint64_t rows[kBlock];
  __shared__ uint32_t scan[kBlock];
  __shared__ uint32_t hist[kSurvBins];
  __shared__ uint32_t s_nsel;
  __shared__ int s_pass_bin[kRefinePasses];
  const uint32_t q = blockIdx.x;
  const int k = (int)kq[q];
  load_query(Q, q, L, qs);
  topk_reset(tk);
  for (int i = threadIdx.x; i < kSurvBins; i += kBlock) hist[i] = 0;
  __syncthreads();
  uint32_t nc = cand_cnt[q];
  // candidates of the approximate stage that found no room: answered again
  const bool cand_dropped = nc > cap;
  if (nc > cap) nc = cap;
  for (uint32_t t0 = 0; t0 < nc; t0 += kBlock) {
    uint32_t i = t0 + threadIdx.x;
    if (i < nc) {
      uint64_t key = cand[(uint64_t)q * cap + i];
      if ((uint32_t)(key >> 32) <= tk.kth_bits) topk_stage(tk, key);
    }
    __syncthreads();
    topk_merge(tk, scan, k);
  }
  uint32_t ns = scnt[q];
  const bool dropped = ns > scap || cand_dropped;
  if (ns > scap) ns = scap;
  const uint64_t *sv = surv + (uint64_t)q * scap;
  const float t0v = thr0[q];
  const float scale = t0v < FLT_MAX && t0v > 0.0f ? (float)kSurvBins / t0v : 0.0f;
  // histogram of the survivors' bounds over [0, threshold]
  for (uint32_t i = threadIdx.x; i < ns; i += kBlock) {
    const float lbv = __uint_as_float((uint32_t)(sv[i] >> 32));
    const int bin = min(kSurvBins - 1, (int)(lbv * scale));
    atomicAdd(&hist[bin], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // last bin of every pass: the pass reaches kRefineFirst * 4^pass
    // survivors counted from the smallest bound (the last pass: the rest)
    uint32_t cum = 0, want = kRefineFirst;
    int b = 0, p = 0;
    for (; b < kSurvBins - 1 && p < kRefinePasses - 1; ++b) {
      cum += hist[b];
      while (p < kRefinePasses - 1 && cum >= want) {
        s_pass_bin[p++] = b;
        want *= 4;
      }
    }
    for (; p < kRefinePasses; ++p) s_pass_bin[p] = kSurvBins - 1;
  }
  __syncthreads();
  uint32_t nref = 0;
  for (int pass = 0; pass < kRefinePasses; ++pass) {
    const int bin_lo = pass == 0 ? -1 : s_pass_bin[pass - 1], bin_hi = s_pass_bin[pass];
    if (bin_hi <= bin_lo) continue;  // uniform: an empty pass
    // the list is read four rounds at a time (four loads in flight per
    // thread): a round's work is short, so one load per round would leave
    // a long list waiting on memory latency
    uint64_t ev[4];
    for (uint32_t t0 = 0; t0 < ns; t0 += kBlock) {
      const uint32_t u = (t0 / kBlock) & 3;
      if (u == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const uint32_t iv = t0 + v * kBlock + threadIdx.x;
          ev[v] = iv < ns ? sv[iv] : 0;
        }
      }
      if (threadIdx.x == 0) s_nsel = 0;
      __syncthreads();
      const uint32_t bnd = tk.kth_bits;
      const uint32_t i = t0 + threadIdx.x;
      bool take = false;
      uint64_t e = 0;
      if (i < ns) {
        e = u == 0 ? ev[0] : u == 1 ? ev[1] : u == 2 ? ev[2] : ev[3];
        const float lbv = __uint_as_float((uint32_t)(e >> 32));
        const int bin = min(kSurvBins - 1, (int)(lbv * scale));
        take = bin > bin_lo && bin <= bin_hi && lb_keep(lbv, bnd);
      }
      const uint64_t m = __ballot(take);
      const int lane = threadIdx.x & 63;
      const int leader = __ffsll((long long)m) - 1;
      uint32_t wb = 0;
      if (m) {
        if (lane == leader) wb = atomicAdd(&s_nsel, (uint32_t)__popcll(m));
        wb = __shfl(wb, leader);
      }
      if (take) rows[wb + __popcll(m & ((1ull << lane) - 1))] = (uint32_t)e;
      __syncthreads();
      refine_rows(X, L, qs, rows, (int)s_nsel, ids, bnd, tk, rs, nref);
      __syncthreads();
      topk_merge(tk, scan, k);
    }
  }
  if ((threadIdx.x & 63) == 0 && nref) atomicAdd(&counters[kCtrRefined], (unsigned long long)nref);
  const uint32_t bq = bound[q];
  const bool uncertified = bq < kInfBits && (tk.cnt < k || tk.kth_bits > bq);
  if (threadIdx.x == 0 && (dropped || uncertified)) {
    atomicAdd(overflow, 1u);
    scnt[q] = 0xFFFFFFFFu;  // tells the caller to answer this query again
  }
  for (uint32_t r = threadIdx.x; r < kmax; r += kBlock) {
    uint64_t id = ~0ull;
    float d = FLT_MAX;
    if ((int)r < tk.cnt) {
      uint64_t key = tk.keys[r];
      d = __uint_as_float((uint32_t)(key >> 32));
      uint32_t id32 = (uint32_t)key;
      id = idmap64 ? idmap64[id32] : (uint64_t)id32;
    }
    out_ids[(uint64_t)q * kmax + r] = id;
    out_d2[(uint64_t)q * kmax + r] = d;
  }
}

void launch_bf16_refine(hipStream_t st, uint32_t nq, const float *X, uint32_t L, const float *Q, const uint32_t *kq,
                        uint32_t kmax, const uint32_t *ids, uint32_t cap, const uint64_t *cand,
                        const uint32_t *cand_cnt, uint32_t scap, const uint64_t *surv, uint32_t *scnt,
                        const float *thr0, const uint32_t *bound, const uint64_t *idmap64, uint64_t *out_ids,
                        float *out_d2, uint32_t *overflow, unsigned long long *counters)
{
  hipLaunchKernelGGL(k_bf16_refine, dim3(nq), dim3(kBlock), L * sizeof(float), st, X, L, Q, kq, kmax, ids, cap, cand,
                     cand_cnt, scap, surv, scnt, thr0, bound, idmap64, out_ids, out_d2, overflow, counters);
}

// reset per-query state: bound = the seed (an upper bound on the k-th
// distance known from an earlier pass) or +inf, no candidates
__global__ void k_reset_queries(uint32_t nq, uint32_t *__restrict__ bound, uint32_t *__restrict__ cand_cnt,
                                const uint32_t *__restrict__ seed)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nq) {
    bound[i] = seed ? seed[i] : kInfBits;
    cand_cnt[i] = 0;
  }
}

__global__ void k_check_k(uint32_t nq, uint32_t kmax, uint32_t *__restrict__ kq, uint32_t *__restrict__ bad)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool out = i < nq && (kq[i] == 0 || kq[i] > kmax);
  if (out) kq[i] = kq[i] == 0 ? 1u : kmax;
  const uint64_t m = __ballot(out);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(bad, (uint32_t)__popcll(m));
}

// small fill kernel (used instead of hipMemsetAsync, which must not be
// recorded into graphs)
__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t n, uint32_t v)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_reset(hipStream_t st, uint32_t nq, uint32_t *bound, uint32_t *cand_cnt, const uint32_t *seed)
{
  hipLaunchKernelGGL(k_reset_queries, dim3((nq + 255) / 256), dim3(256), 0, st, nq, bound, cand_cnt, seed);
}

__global__ void k_stamp(unsigned long long *__restrict__ t)
{
  if (threadIdx.x == 0) *t = (unsigned long long)wall_clock64();
}

void launch_stamp(hipStream_t st, unsigned long long *t)
{
  hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, st, t);
}

void launch_check_k(hipStream_t st, uint32_t nq, uint32_t kmax, uint32_t *kq, uint32_t *bad)
{
  hipLaunchKernelGGL(k_check_k, dim3((nq + 255) / 256), dim3(256), 0, st, nq, kmax, kq, bad);
}

void launch_fill(hipStream_t st, uint32_t *p, uint32_t n, uint32_t v)
{
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}

void launch_prep(hipStream_t st, const float *Q, uint32_t nq, uint32_t L, float *paa, double *psum, double *qn2)
{
  hipLaunchKernelGGL(k_prep, dim3(nq), dim3(64), 0, st, Q, L, paa, psum, qn2);
}

void launch_eapca(hipStream_t st, uint32_t nq, uint32_t nleaf, uint32_t L, const uint32_t *seg_off,
                  const int16_t *seg_end, const float4 *ind, const double *psum, float *lb)
{
  size_t lds = kEapcaQueries * 2 * (L + 1) * sizeof(double);
  if (lds > kEapcaLdsBytes) lds = 0;
  hipLaunchKernelGGL(k_eapca, dim3((nleaf + kBlock - 1) / kBlock, (nq + kEapcaQueries - 1) / kEapcaQueries),
                     dim3(kBlock), lds, st, nleaf, L, nq, seg_off, seg_end, ind, psum, lb);
}

void launch_best_leaf(hipStream_t st, uint32_t nq, uint32_t nleaf, const float *lb, uint32_t *best)
{
  hipLaunchKernelGGL(k_best_leaf, dim3(nq), dim3(kBlock), 0, st, nleaf, lb, best);
}

void launch_scan_best_bf16(hipStream_t st, uint32_t nq, const float *X, const uint16_t *Xb, const float *xn2,
                           uint32_t max_leaf, uint32_t L, const float *Q, const uint32_t *kq, const uint8_t *sax,
                           const float *paa, const uint32_t *best,
                           const uint64_t *leaf_pos, const uint32_t *leaf_size, const uint32_t *ids, uint32_t cap,
                           uint64_t *cand, uint32_t *cand_cnt, uint32_t *bound, unsigned long long *counters)
{
  const size_t dyn = (size_t)L * sizeof(float) + std::max<size_t>((size_t)max_leaf * sizeof(uint32_t), sizeof(RefineShared));
  if (dyn > kApproxMaxDynLds) {
    // leaves too large for the ranking keys in LDS: every row exactly
    launch_scan_best(st, nq, X, L, Q, kq, best, leaf_pos, leaf_size, ids, cap, cand, cand_cnt, bound, counters);
    return;
  }
  // the best leaf of every query first, then the others with the rows whose
  // iSAX bound is within the k-th distance that leaf gave
  hipLaunchKernelGGL((k_scan_best_bf16<false>), dim3(nq, 1), dim3(kBlock), dyn, st, X, Xb, xn2, L, Q, kq, sax, paa,
                     0u, best, leaf_pos, leaf_size, ids, cap, cand, cand_cnt, bound, counters);
  if (kBestLeaves > 1)
    hipLaunchKernelGGL((k_scan_best_bf16<true>), dim3(nq, kBestLeaves - 1), dim3(kBlock), dyn, st, X, Xb, xn2, L, Q,
                       kq, sax, paa, 1u, best, leaf_pos, leaf_size, ids, cap, cand, cand_cnt, bound, counters);
}

void launch_scan_best(hipStream_t st, uint32_t nq, const float *X, uint32_t L, const float *Q,
                      const uint32_t *kq, const uint32_t *best, const uint64_t *leaf_pos,
                      const uint32_t *leaf_size, const uint32_t *ids, uint32_t cap, uint64_t *cand,
                      uint32_t *cand_cnt, uint32_t *bound, unsigned long long *counters)
{
  hipLaunchKernelGGL(k_scan_best, dim3(nq, kBestLeaves), dim3(kBlock), L * sizeof(float), st, X, L, Q, kq, best,
                     leaf_pos, leaf_size, ids, cap, cand, cand_cnt, bound, counters);
}

void launch_collect(hipStream_t st, uint32_t nq, uint32_t nleaf, const float *lb, const uint32_t *best,
                    const uint32_t *bound, uint32_t *keys, uint32_t *vals, uint32_t *nwork)
{
  hipLaunchKernelGGL(k_collect, dim3((nleaf + kBlock - 1) / kBlock, nq), dim3(kBlock), 0, st, nq, nleaf, lb, best,
                     bound, keys, vals, nwork);
}

size_t sort_work_temp_bytes(uint32_t n)
{
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                           (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, 32);
  return bytes;
}

hipError_t launch_sort_work(hipStream_t st, void *tmp, size_t tmp_bytes, const uint32_t *keys, uint32_t *keys_out,
                            const uint32_t *vals, uint32_t *vals_out, uint32_t n)
{
  return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys_out, vals, vals_out, (int)n, 0, 32, st);
}

void launch_filter_refine(hipStream_t st, uint32_t grid, const float *X, const uint8_t *sax, uint32_t L,
                          const float *Q, const float *paa, const uint32_t *kq, const uint64_t *leaf_pos,
                          const uint32_t *leaf_size, const uint32_t *ids, uint32_t nleaf, const float *lb,
                          const uint32_t *work, const uint32_t *nwork, uint32_t *work_next, uint32_t cap,
                          uint64_t *cand, uint32_t *cand_cnt, uint32_t *bound, unsigned long long *counters)
{
  hipLaunchKernelGGL(k_filter_refine, dim3(grid), dim3(kBlock), L * sizeof(float), st, X, sax, L, Q, paa, kq,
                     leaf_pos, leaf_size, ids, nleaf, lb, work, nwork, work_next, cap, cand, cand_cnt, bound,
                     counters);
}

void launch_scan(hipStream_t st, uint32_t grid_x, const float *X, uint64_t n, uint32_t L, const float *Q,
                 uint32_t nq, const uint32_t *kq, const uint32_t *ids, uint32_t cap, uint64_t *cand,
                 uint32_t *cand_cnt, uint32_t *bound)
{
  // queries per pass: up to kScanQ (a lone query scans alone)
  const uint32_t qn = scan_group_size(nq, L);
  const uint32_t groups = (nq + qn - 1) / qn;
  const size_t lds = (size_t)qn * L * sizeof(float);
  switch (qn) {
  case 1: hipLaunchKernelGGL(k_scan<1>, dim3(grid_x, groups), dim3(kBlock), lds, st, X, n, L, Q, nq, kq, ids, cap, cand, cand_cnt, bound); break;
  case 2: hipLaunchKernelGGL(k_scan<2>, dim3(grid_x, groups), dim3(kBlock), lds, st, X, n, L, Q, nq, kq, ids, cap, cand, cand_cnt, bound); break;
  default: hipLaunchKernelGGL(k_scan<kScanQ>, dim3(grid_x, groups), dim3(kBlock), lds, st, X, n, L, Q, nq, kq, ids, cap, cand, cand_cnt, bound); break;
  }
}

// for the current device: the query kernels may take kQueryLdsMax bytes of
// dynamic LDS (a series of 16384 floats is 64 KB)
void set_lds_limits()
{
  const void *fns[] = {reinterpret_cast<const void *>(k_scan_best), reinterpret_cast<const void *>(k_filter_refine),
                       reinterpret_cast<const void *>(k_bf16_refine),
                       reinterpret_cast<const void *>(k_scan<1>), reinterpret_cast<const void *>(k_scan<2>),
                       reinterpret_cast<const void *>(k_scan<kScanQ>)};
  for (const void *f : fns)
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kQueryLdsMax);
  (void)hipGetLastError();
}

void launch_select(hipStream_t st, uint32_t nq, uint32_t kmax, const uint32_t *kq, uint32_t cap,
                   const uint64_t *cand, const uint32_t *cand_cnt, const uint64_t *idmap64, uint64_t *out_ids,
                   float *out_d2, uint32_t *overflow)
{
  hipLaunchKernelGGL(k_select, dim3(nq), dim3(kBlock), 0, st, kmax, kq, cap, cand, cand_cnt, idmap64, out_ids,
                     out_d2, overflow);
}


// ---------------------------------------------------------------------------
// approximate modes: the reference's best-first traversals
// ---------------------------------------------------------------------------
// The epsilon search of exact_de_knn_search_psq_lf_sims_adaptive
// (hercules_query_engine.c:706-1083) and exact_ng_knn_search (:2623-2799)
// decide which leaves they examine from the exact FP32 node bounds, a
// binary heap's pop order and the k-th distance of the moment, so a block
// per query replays them: lane 0 runs the heap, the block scans each leaf
// the heap hands out.  Node bounds are the reference's own bits.

__device__ __forceinline__ float tree8f(const float *v)
{
  return ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}

// calc_mean_stdev_SIMD (calc_utils.c:48-112): 8-chunks from the segment
// start, the last zero-padded, FP32 sums; sd is NaN when the FP32 variance
// is negative (every test it enters is then false, as in the reference)
__device__ void ref_seg_stats(const float *x, int s, int e, float &mu, float &sd)
{
  const int n = e - s;
  float sx = 0.0f, sq = 0.0f;
  for (int b = s; b < e; b += 8) {
    float c[8], c2[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      c[j] = b + j < e ? x[b + j] : 0.0f;
      c2[j] = c[j] * c[j];
    }
    sx = sx + tree8f(c);
    sq = sq + tree8f(c2);
  }
  mu = sx / (float)n;
  sq = sq - (sx * sx) / (float)n;
  sd = (float)sqrt((double)(sq / (float)n));
}

// calculate_node_min_distance (hercules_node.c:879-924) as GCC compiles it
// for an FMA target: three contractions (both temp_dist += temp * temp and
// sum += temp_dist * len), every other operation rounded on its own
__device__ float lb_eapca_ref(const float *qs, const int16_t *se, const float4 *ind, int nseg)
{
  float sum = 0.0f;
  int start = 0;
  for (int i = 0; i < nseg; ++i) {
    const int end = se[i];
    float mu, sd;
    ref_seg_stats(qs, start, end, mu, sd);
    const float4 g = ind[i];  // {max mean, min mean, max sd, min sd}
    float td = 0.0f;
    if ((sd - g.z) * (sd - g.w) > 0.0f) {
      float t = fminf(fabsf(sd - g.z), fabsf(sd - g.w));
      td = fmaf(t, t, td);
    }
    if ((mu - g.x) * (mu - g.y) > 0.0f) {
      float t = fminf(fabsf(mu - g.x), fabsf(mu - g.y));
      td = fmaf(t, t, td);
    }
    sum = fmaf(td, (float)(end - start), sum);
    start = end;
  }
  return sum;
}

// exact node bound of every (query, node): one lane per node
__global__ void __launch_bounds__(kBlock) k_node_lb(uint32_t nnodes, uint32_t L, const float *__restrict__ Q,
                                                    const uint2 *__restrict__ nd_seg,
                                                    const int16_t *__restrict__ seg_end,
                                                    const float4 *__restrict__ ind, float *__restrict__ nlb)
{
  extern __shared__ float qs[];
  const uint32_t q = blockIdx.y;
  load_query(Q, q, L, qs);
  __syncthreads();
  const uint32_t node = blockIdx.x * kBlock + threadIdx.x;
  if (node >= nnodes) return;
  const uint2 s = nd_seg[node];
  nlb[(uint64_t)q * nnodes + node] = s.y ? lb_eapca_ref(qs, seg_end + s.x, ind + s.x, (int)s.y) : 0.0f;
}

// every (query, leaf) work item absent: key all ones, value its index
__global__ void k_mode_keys(uint32_t pairs, uint32_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < pairs) {
    keys[i] = 0xFFFFFFFFu;
    vals[i] = i;
  }
}

// pqueue.c:63-200 with cmp_pri = (next > curr): a min-heap in d[1, size)
// ordered by the bound alone (entries: bound bits << 32 | node; bounds are
// non-negative, so their bits order like their values)
__device__ __forceinline__ uint32_t hpri(uint64_t e) { return (uint32_t)(e >> 32); }

__device__ void heap_push(uint64_t *d, uint32_t &size, float pri, uint32_t node)
{
  uint32_t i = size++;
  const uint64_t mv = ((uint64_t)__float_as_uint(pri) << 32) | node;
  while (i > 1) {
    const uint64_t p = d[i >> 1];
    if (!(hpri(p) > hpri(mv))) break;
    d[i] = p;
    i >>= 1;
  }
  d[i] = mv;
}

__device__ bool heap_pop(uint64_t *d, uint32_t &size, uint64_t &out)
{
  if (size == 1) return false;
  out = d[1];
  const uint64_t mv = d[--size];
  uint32_t i = 1;
  for (;;) {
    uint32_t c = i << 1;
    if (c >= size) break;
    if (c + 1 < size && hpri(d[c]) > hpri(d[c + 1])) ++c;
    if (!(hpri(mv) > hpri(d[c]))) break;
    d[i] = d[c];
    i = c;
  }
  d[i] = mv;
  return true;
}

// rows [beg, end) of a leaf into the block's running answer: one row per
// lane, streamed in 128-byte steps and folded in the reference's order
// (ts.c:83-109), abandoned once the partial sum exceeds the k-th distance
// (partial sums only grow, so no row within it is lost).  Passing rows are
// staged and merged only when the staging area would overflow, and once at
// the end: the answer after the call is the same, with far fewer merges.
__device__ void scan_rows(const float *__restrict__ X, uint32_t L, const float *qs, uint64_t pos, uint32_t beg,
                          uint32_t end, const uint32_t *__restrict__ ids, int k, TopK &tk, uint32_t *scan,
                          uint32_t &nref)
{
  __shared__ uint32_t s_pass;
  const int nch = (int)(L >> 3);
  const float4 *q4 = reinterpret_cast<const float4 *>(qs);
  for (uint32_t t0 = beg; t0 < end; t0 += kBlock) {
    if (threadIdx.x == 0) s_pass = 0;
    __syncthreads();
    const uint32_t bnd = tk.kth_bits;
    const uint32_t r = t0 + threadIdx.x;
    bool pass = false;
    float acc = 0.0f;
    uint64_t row = 0;
    if (r < end) {
      row = pos + r;
      const float4 *xp = reinterpret_cast<const float4 *>(X + row * L);
      bool alive = true;
      int c = 0;
      for (; c + 4 <= nch && alive; c += 4) {
        float4 xv[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) xv[i] = ld_stream(reinterpret_cast<const float *>(xp + 2 * c + i));
#pragma unroll
        for (int sub = 0; sub < 4; ++sub)
          acc = acc + ed_chunk(q4[2 * (c + sub)], q4[2 * (c + sub) + 1], xv[2 * sub], xv[2 * sub + 1]);
        alive = __float_as_uint(acc) <= bnd;
      }
      for (; c < nch && alive; ++c) {
        const float4 xa = ld_stream(reinterpret_cast<const float *>(xp + 2 * c));
        const float4 xb = ld_stream(reinterpret_cast<const float *>(xp + 2 * c + 1));
        acc = acc + ed_chunk(q4[2 * c], q4[2 * c + 1], xa, xb);
        alive = __float_as_uint(acc) <= bnd;
      }
      pass = alive;
    }
    // rows examined: lane 0 of every wave counts its wave's (its nref is
    // the one reported)
    const uint64_t live = __ballot(r < end);
    if ((threadIdx.x & 63) == 0) nref += (uint32_t)__popcll(live);
    const uint64_t bal = __ballot(pass);
    if ((threadIdx.x & 63) == 0 && bal) atomicAdd(&s_pass, (uint32_t)__popcll(bal));
    __syncthreads();
    if (tk.nstage + (int)s_pass > kBlock) topk_merge(tk, scan, k);  // uniform
    if (pass) topk_stage(tk, make_key(acc, ids ? ids[row] : (uint32_t)row));
    __syncthreads();
  }
  topk_merge(tk, scan, k);
}

// one query per block.  ng (nprobes > 0): the leaf the split policies route
// the query to, then best-first probes (exact_ng_knn_search).  epsilon:
// approximate_greedy_knn_search's best-first leaves (up to the budget), the
// candidate leaves whose bound passes kth/(1 + epsilon), and the branch the
// reference takes for them (:893-978): sequential in file order against the
// k-th distance its loop variable holds (exact_de_knn_search_noapprox,
// :2450-2507), or every candidate refined exactly -- then handed to the
// work list of k_filter_refine.  info per query: {leaves scanned first,
// candidate leaves, branch (0 none, 1 sequential, 2 refined), candidate
// leaves scanned}.
__global__ void __launch_bounds__(kBlock) k_mode_search(
    const float *__restrict__ X, const uint8_t *__restrict__ sax, uint32_t L, const float *__restrict__ Q,
    const float *__restrict__ paa, const uint32_t *__restrict__ kq, uint32_t nnodes,
    const int4 *__restrict__ nd_link, const int4 *__restrict__ nd_split, const float *__restrict__ nlb,
    uint32_t nleaf, const uint64_t *__restrict__ leaf_pos, const uint32_t *__restrict__ leaf_size,
    const int32_t *__restrict__ leaf_node, const uint32_t *__restrict__ ids, const float *__restrict__ lb,
    uint64_t *__restrict__ heap_g, int heap_in_lds, uint32_t *__restrict__ bfs, ModeParams mp, uint32_t cap,
    uint64_t *__restrict__ cand,
    uint32_t *__restrict__ cand_cnt, uint32_t *__restrict__ bound, uint32_t *__restrict__ keys,
    uint32_t *__restrict__ nwork, uint32_t *__restrict__ info, unsigned long long *__restrict__ counters,
    uint32_t *__restrict__ mstate, uint64_t *__restrict__ vtop, uint32_t vcap)
{
  extern __shared__ __align__(16) unsigned char mode_lds[];
  float *qs = reinterpret_cast<float *>(mode_lds);
  const uint32_t q = blockIdx.x;
  uint64_t *hp = heap_in_lds ? reinterpret_cast<uint64_t *>(mode_lds + ((L * 4 + 15) & ~15u))
                             : heap_g + (uint64_t)q * (nnodes + 1);
  __shared__ TopK tk;
  __shared__ uint32_t scan[kBlock];
  __shared__ float bp[256];
  __shared__ float spaa[kPaaSegments];
  __shared__ int32_t s_node;
  const int k = (int)kq[q];
  const bool ng = mp.nprobes > 0;
  const float *qnlb = nlb + (uint64_t)q * nnodes;
  uint32_t *qkeys = keys + (uint64_t)q * nleaf;
  load_query(Q, q, L, qs);
  for (int i = threadIdx.x; i < 256; i += kBlock) bp[i] = c_bp[i];
  if (threadIdx.x < kPaaSegments) spaa[threadIdx.x] = paa[q * kPaaSegments + threadIdx.x];
  topk_reset(tk);
  __syncthreads();
  uint32_t nref = 0;
  unsigned long long first_rows = 0;
  // lane 0's state
  uint32_t hsize = 1, nvis = 0, ncand = 0, nempty = 0;
  float best = FLT_MAX;
  int32_t best_leaf = -1;
  if (ng) {
    if (threadIdx.x == 0) {
      int32_t i = 0;
      for (int4 ln = nd_link[0]; !ln.w; ln = nd_link[i]) {
        const int4 sp = nd_split[i];  // {from, to, indicator (0 mean, 1 sd), value bits}
        float mu, sd;
        ref_seg_stats(qs, sp.x, sp.y, mu, sd);
        i = ((sp.z & 1) ? sd : mu) < __int_as_float(sp.w) ? ln.x : ln.y;
      }
      s_node = i;
    }
    __syncthreads();
    const int32_t lf = nd_link[s_node].z;
    if (lf >= 0) {
      scan_rows(X, L, qs, leaf_pos[lf], 0, leaf_size[lf], ids, k, tk, scan, nref);
      first_rows += leaf_size[lf];
    }
    if (threadIdx.x == 0) {
      nvis = 1;
      if (tk.cnt > 0) {
        best = __uint_as_float((uint32_t)(tk.keys[0] >> 32));
        best_leaf = s_node;
      }
      heap_push(hp, hsize, qnlb[0], 0);
    }
  } else if (threadIdx.x == 0) {
    heap_push(hp, hsize, qnlb[0], 0);
  }
  // best-first leaves: the greedy approximate phase or the ng probes
  const uint32_t budget = mp.budget ? mp.budget : 1u;
  for (;;) {
    if (threadIdx.x == 0) {
      int32_t next = -1;
      if (ng || nvis < budget) {
        uint64_t it;
        while (heap_pop(hp, hsize, it)) {
          if (ng && nvis >= mp.nprobes) break;  // popped, then the probe count ends the loop
          const float pri = __uint_as_float(hpri(it));
          const uint32_t node = (uint32_t)it;
          const float kth = __uint_as_float(tk.kth_bits);
          if (pri > (ng ? best : kth)) break;   // the popped item is dropped
          const int4 ln = nd_link[node];
          if (ln.w) {
            next = (int32_t)node;
            break;
          }
          for (int c = 0; c < 2; ++c) {
            const int32_t ch = c ? ln.y : ln.x;
            const float b = qnlb[ch];
            if (b < kth && ch != best_leaf) heap_push(hp, hsize, b, (uint32_t)ch);
          }
        }
      }
      s_node = next;
    }
    __syncthreads();
    const int32_t node = s_node;
    if (node < 0) break;
    const int32_t lf = nd_link[node].z;
    if (lf >= 0) {
      scan_rows(X, L, qs, leaf_pos[lf], 0, leaf_size[lf], ids, k, tk, scan, nref);
      first_rows += leaf_size[lf];
    }
    if (threadIdx.x == 0) ++nvis;
    __syncthreads();
  }
  uint32_t branch = 0, nscanned = 0;
  if (!ng) {
    __shared__ float s_kth0;
    __shared__ int s_branch;
    __shared__ uint32_t s_hsize, s_ncand, s_nempty, s_nnext;
    __shared__ unsigned long long s_crows;
    if (threadIdx.x == 0) {
      s_kth0 = __uint_as_float(tk.kth_bits);
      s_hsize = hsize;
      s_ncand = 0;
      s_nempty = 0;
      s_nnext = 0;
      s_crows = 0;
    }
    __syncthreads();
    // candidate leaves (:836-891): every queued node within kth0 / (1 +
    // epsilon) and all it reaches through children below that bound.  The
    // bound no longer moves, so the queue's order decides only when a leaf
    // is found, not which: the block walks the tree level by level.
    const float thr = s_kth0 / (1.0f + mp.epsilon);
    uint32_t *cur = bfs + (uint64_t)q * 2 * (nnodes + 1), *nxt = cur + (nnodes + 1);
    auto visit = [&](uint32_t node) __attribute__((always_inline)) {
      const int4 ln = nd_link[node];
      if (ln.w) {
        atomicAdd(&s_ncand, 1u);
        if (ln.z >= 0) {
          qkeys[ln.z] = __float_as_uint(lb[(uint64_t)q * nleaf + ln.z]);
          atomicAdd(&s_crows, (unsigned long long)leaf_size[ln.z]);
        } else {
          atomicAdd(&s_nempty, 1u);
        }
      } else {
        if (qnlb[ln.x] < thr) nxt[atomicAdd(&s_nnext, 1u)] = (uint32_t)ln.x;
        if (qnlb[ln.y] < thr) nxt[atomicAdd(&s_nnext, 1u)] = (uint32_t)ln.y;
      }
    };
    for (uint32_t i = 1 + threadIdx.x; i < s_hsize; i += kBlock) {
      const uint64_t e = hp[i];
      if (__uint_as_float(hpri(e)) <= thr) visit((uint32_t)e);
    }
    __syncthreads();
    for (;;) {
      const uint32_t n = s_nnext;
      __syncthreads();
      if (n == 0) break;
      if (threadIdx.x == 0) s_nnext = 0;
      uint32_t *t = cur;
      cur = nxt;
      nxt = t;
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < n; i += kBlock) visit(cur[i]);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      ncand = s_ncand;
      nempty = s_nempty;
      int br = 0;
      if (ncand) {
        const float flp = (float)(1.0 - (double)ncand / (double)(float)mp.nleaves);
        br = flp < mp.eapca_th ? 1 : 3;  // 3: decided by the iSAX count
        // the count is at most every candidate row, and 1 - count / n only
        // falls as the count grows (each rounding is monotone): when even
        // all of them keep it at or above the threshold, so does the count
        if (br == 3 && !((float)(1.0 - (double)s_crows / (double)(float)mp.nseries) < mp.sax_th)) br = 2;
      }
      s_branch = br;
    }
    __syncthreads();
    const int br = s_branch;
    branch = (uint32_t)br;
    if (br == 2) {
      nscanned = ncand - nempty;
      if (threadIdx.x == 0 && nscanned) atomicAdd(nwork, nscanned);
    }
  }
  // the sequential branch, and the iSAX count that may choose it, run in
  // k_mode_count / k_mode_finish: hand them this query's state and answer
  const bool defer = branch == 1 || branch == 3;
  if (threadIdx.x == 0) {
    uint32_t *ms = mstate + (uint64_t)q * kModeState;
    ms[3] = defer ? branch : 0u;
    if (defer) {
      ms[0] = tk.kth_bits;  // kth0: the k-th distance after the approximate phase
      ms[1] = ncand;
      ms[2] = nempty;
      ms[4] = (uint32_t)tk.cnt;
    }
  }
  if (defer) {
    for (int i = threadIdx.x; i < tk.cnt; i += kBlock) vtop[(uint64_t)q * vcap + i] = tk.keys[i];
  } else {
    flush_topk(tk, k, q, cap, cand, cand_cnt, bound);
  }
  if (threadIdx.x == 0) {
    info[4 * q + 0] = nvis;
    info[4 * q + 1] = ncand;
    info[4 * q + 2] = branch;
    info[4 * q + 3] = nscanned;
    atomicAdd(&counters[kCtrApprox], first_rows);
    atomicAdd(&counters[kCtrModeFirst], (unsigned long long)nvis);
    atomicAdd(&counters[kCtrModeCand], (unsigned long long)ncand);
    if (!defer && nscanned) atomicAdd(&counters[kCtrModeScanned], (unsigned long long)nscanned);
  }
  if ((threadIdx.x & 63) == 0 && nref) atomicAdd(&counters[kCtrRefined], (unsigned long long)nref);
}

// The iSAX count of the epsilon search's queries whose branch it decides
// (:2171-2205): the rows of their candidate leaves whose bound is within
// kth0, every (query, 256-leaf chunk) a block, one wave per leaf.
__global__ void __launch_bounds__(kBlock) k_mode_count(
    const uint8_t *__restrict__ sax, uint32_t L, const float *__restrict__ paa, uint32_t nnodes,
    const float *__restrict__ nlb, uint32_t nleaf, const uint64_t *__restrict__ leaf_pos,
    const uint32_t *__restrict__ leaf_size, const int32_t *__restrict__ leaf_node,
    const uint32_t *__restrict__ keys, const uint32_t *__restrict__ mstate, unsigned long long *__restrict__ lab)
{
  const uint32_t q = blockIdx.y;
  const uint32_t *ms = mstate + (uint64_t)q * kModeState;
  if (ms[3] != 3u) return;  // uniform
  __shared__ float bp[256];
  __shared__ float spaa[kPaaSegments];
  __shared__ uint64_t s_lpos[kBlock];
  __shared__ uint32_t s_lsize[kBlock];
  __shared__ int s_n;
  __shared__ unsigned long long s_cnt;
  const float kth0 = __uint_as_float(ms[0]);
  for (int i = threadIdx.x; i < 256; i += kBlock) bp[i] = c_bp[i];
  if (threadIdx.x < kPaaSegments) spaa[threadIdx.x] = paa[q * kPaaSegments + threadIdx.x];
  if (threadIdx.x == 0) {
    s_n = 0;
    s_cnt = 0;
  }
  __syncthreads();
  const uint32_t l = blockIdx.x * kBlock + threadIdx.x;
  if (l < nleaf && keys[(uint64_t)q * nleaf + l] != 0xFFFFFFFFu &&
      nlb[(uint64_t)q * nnodes + leaf_node[l]] <= kth0) {
    const int j = atomicAdd(&s_n, 1);
    s_lpos[j] = leaf_pos[l];
    s_lsize[j] = leaf_size[l];
  }
  __syncthreads();
  const int n = s_n;
  const float ratio = (float)(L / kPaaSegments);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned long long mine = 0;
  for (int j = wave; j < n; j += kBlock / 64) {
    const uint64_t pos = s_lpos[j];
    const uint32_t size = s_lsize[j];
    for (uint32_t r = lane; r < size; r += 8 * 64) {
      uint4 w[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t rr = r + u * 64;
        w[u] = rr < size ? *reinterpret_cast<const uint4 *>(sax + (pos + rr) * kPaaSegments) : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (r + u * 64 < size) mine += lb_sax_word(spaa, w[u], bp, ratio) <= kth0 ? 1ull : 0ull;
    }
  }
  if (mine) atomicAdd(&s_cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0 && s_cnt) atomicAdd(&lab[q], s_cnt);
}

// The epsilon search's deferred queries: the branch the iSAX count chose
// (:967-978), then either every candidate to the work list, or the
// candidates scanned in file order against the k-th distance the
// reference's loop variable holds (exact_de_knn_search_noapprox,
// :2450-2507); the answer so far comes back from k_mode_search's hand-over.
__global__ void __launch_bounds__(kBlock) k_mode_finish(
    const float *__restrict__ X, uint32_t L, const float *__restrict__ Q, const uint32_t *__restrict__ kq,
    uint32_t nnodes, const float *__restrict__ nlb, uint32_t nleaf, const uint64_t *__restrict__ leaf_pos,
    const uint32_t *__restrict__ leaf_size, const int32_t *__restrict__ leaf_node,
    const uint32_t *__restrict__ ids, ModeParams mp, uint32_t cap, uint64_t *__restrict__ cand,
    uint32_t *__restrict__ cand_cnt, uint32_t *__restrict__ bound, uint32_t *__restrict__ keys,
    uint32_t *__restrict__ nwork, uint32_t *__restrict__ info, unsigned long long *__restrict__ counters,
    const uint32_t *__restrict__ mstate, const uint64_t *__restrict__ vtop, uint32_t vcap,
    const unsigned long long *__restrict__ lab)
{
  const uint32_t q = blockIdx.x;
  const uint32_t *ms = mstate + (uint64_t)q * kModeState;
  uint32_t br = ms[3];
  if (br == 0u) return;  // uniform: answered by k_mode_search
  extern __shared__ float qs[];
  __shared__ TopK tk;
  __shared__ uint32_t scan[kBlock];
  __shared__ uint32_t s_list[kBlock];
  __shared__ float s_kthb;
  __shared__ int s_go;
  const int k = (int)kq[q];
  const uint32_t ncand = ms[1], nempty = ms[2];
  if (br == 3u) {
    const float slp = (float)(1.0 - (double)lab[q] / (double)(float)mp.nseries);
    br = slp < mp.sax_th ? 1u : 2u;
  }
  load_query(Q, q, L, qs);
  // the answer after the approximate phase, as k_mode_search left it
  const int cnt = (int)ms[4];
  for (int i = threadIdx.x; i < cnt; i += kBlock) tk.keys[i] = vtop[(uint64_t)q * vcap + i];
  if (threadIdx.x == 0) {
    tk.cnt = cnt;
    tk.nstage = 0;
    tk.kth_bits = cnt == k ? (uint32_t)(tk.keys[k - 1] >> 32) : kInfBits;
  }
  __syncthreads();
  uint32_t nref = 0, nscanned = 0;
  uint32_t *qkeys = keys + (uint64_t)q * nleaf;
  const float *qnlb = nlb + (uint64_t)q * nnodes;
  if (br == 2u) {
    nscanned = ncand - nempty;
    if (threadIdx.x == 0 && nscanned) atomicAdd(nwork, nscanned);
  } else {
    if (threadIdx.x == 0) s_kthb = __uint_as_float(tk.kth_bits);
    for (uint32_t l0 = 0; l0 < nleaf; l0 += kBlock) {
      const uint32_t l = l0 + threadIdx.x;
      const bool is_c = l < nleaf && qkeys[l] != 0xFFFFFFFFu;
      __syncthreads();
      // ordered compaction of this chunk's candidates
      scan[threadIdx.x] = is_c ? 1u : 0u;
      __syncthreads();
      for (int off = 1; off < kBlock; off <<= 1) {
        uint32_t v = threadIdx.x >= (uint32_t)off ? scan[threadIdx.x - off] : 0u;
        __syncthreads();
        scan[threadIdx.x] += v;
        __syncthreads();
      }
      if (is_c) {
        s_list[scan[threadIdx.x] - 1] = l;
        qkeys[l] = 0xFFFFFFFFu;  // handled here, not by the work list
      }
      const int nl = (int)scan[kBlock - 1];
      __syncthreads();
      for (int j = 0; j < nl; ++j) {
        const uint32_t lf = s_list[j];
        if (threadIdx.x == 0) s_go = qnlb[leaf_node[lf]] <= s_kthb / (1.0f + mp.epsilon);
        __syncthreads();
        const uint32_t size = leaf_size[lf];
        if (s_go && size > 0) {
          const uint64_t pos = leaf_pos[lf];
          scan_rows(X, L, qs, pos, 0, size - 1, ids, k, tk, scan, nref);
          if (threadIdx.x == 0) s_kthb = __uint_as_float(tk.kth_bits);  // read before the last row
          scan_rows(X, L, qs, pos, size - 1, size, ids, k, tk, scan, nref);
          ++nscanned;
        }
        __syncthreads();
      }
    }
  }
  flush_topk(tk, k, q, cap, cand, cand_cnt, bound);
  if (threadIdx.x == 0) {
    info[4 * q + 2] = br;
    info[4 * q + 3] = nscanned;
    if (nscanned) atomicAdd(&counters[kCtrModeScanned], (unsigned long long)nscanned);
    if (br == 1u) atomicAdd(&counters[kCtrModeSeq], 1ull);
  }
  if ((threadIdx.x & 63) == 0 && nref) atomicAdd(&counters[kCtrRefined], (unsigned long long)nref);
}

void launch_mode_count(hipStream_t st, uint32_t nq, const uint8_t *sax, uint32_t L, const float *paa, uint32_t nnodes,
                       const float *nlb, uint32_t nleaf, const uint64_t *leaf_pos, const uint32_t *leaf_size,
                       const int32_t *leaf_node, const uint32_t *keys, const uint32_t *mstate, unsigned long long *lab)
{
  hipLaunchKernelGGL(k_mode_count, dim3((nleaf + kBlock - 1) / kBlock, nq), dim3(kBlock), 0, st, sax, L, paa, nnodes,
                     nlb, nleaf, leaf_pos, leaf_size, leaf_node, keys, mstate, lab);
}

void launch_mode_finish(hipStream_t st, uint32_t nq, const float *X, uint32_t L, const float *Q, const uint32_t *kq,
                        uint32_t nnodes, const float *nlb, uint32_t nleaf, const uint64_t *leaf_pos,
                        const uint32_t *leaf_size, const int32_t *leaf_node, const uint32_t *ids, const ModeParams &mp,
                        uint32_t cap, uint64_t *cand, uint32_t *cand_cnt, uint32_t *bound, uint32_t *keys,
                        uint32_t *nwork, uint32_t *info, unsigned long long *counters, const uint32_t *mstate,
                        const uint64_t *vtop, uint32_t vcap, const unsigned long long *lab)
{
  hipLaunchKernelGGL(k_mode_finish, dim3(nq), dim3(kBlock), L * sizeof(float), st, X, L, Q, kq, nnodes, nlb, nleaf,
                     leaf_pos, leaf_size, leaf_node, ids, mp, cap, cand, cand_cnt, bound, keys, nwork, info, counters,
                     mstate, vtop, vcap, lab);
}

void launch_node_lb(hipStream_t st, uint32_t nq, uint32_t nnodes, uint32_t L, const float *Q, const uint2 *nd_seg,
                    const int16_t *seg_end, const float4 *ind, float *nlb)
{
  hipLaunchKernelGGL(k_node_lb, dim3((nnodes + kBlock - 1) / kBlock, nq), dim3(kBlock), L * sizeof(float), st,
                     nnodes, L, Q, nd_seg, seg_end, ind, nlb);
}

void launch_mode_keys(hipStream_t st, uint32_t pairs, uint32_t *keys, uint32_t *vals)
{
  hipLaunchKernelGGL(k_mode_keys, dim3((pairs + 255) / 256), dim3(256), 0, st, pairs, keys, vals);
}

// The queue only ever holds roots of disjoint subtrees, so it never has more
// entries than the tree has leaves ((nnodes + 1) / 2 in a full binary tree).
static size_t mode_heap_bytes(uint32_t nnodes) { return ((size_t)(nnodes + 1) / 2 + 2) * 8; }
static size_t mode_qs_bytes(uint32_t L) { return ((size_t)L * 4 + 15) & ~(size_t)15; }

bool mode_heap_in_lds(uint32_t L, uint32_t nnodes)
{
  return mode_qs_bytes(L) + mode_heap_bytes(nnodes) <= kModeLdsMax;
}

// the mode kernels that hold the query in LDS (as set_lds_limits)
void set_mode_lds_limits()
{
  const void *fns[] = {reinterpret_cast<const void *>(k_node_lb), reinterpret_cast<const void *>(k_mode_finish)};
  for (const void *f : fns)
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kQueryLdsMax);
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_mode_search), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kModeLdsMax);
  (void)hipGetLastError();
}

void launch_mode_search(hipStream_t st, uint32_t nq, const float *X, const uint8_t *sax, uint32_t L, const float *Q,
                        const float *paa, const uint32_t *kq, uint32_t nnodes, const int4 *nd_link,
                        const int4 *nd_split, const float *nlb, uint32_t nleaf, const uint64_t *leaf_pos,
                        const uint32_t *leaf_size, const int32_t *leaf_node, const uint32_t *ids, const float *lb,
                        uint64_t *heap_g, uint32_t *bfs, const ModeParams &mp, uint32_t cap, uint64_t *cand,
                        uint32_t *cand_cnt, uint32_t *bound, uint32_t *keys, uint32_t *nwork, uint32_t *info,
                        unsigned long long *counters, uint32_t *mstate, uint64_t *vtop, uint32_t vcap)
{
  const bool in_lds = mode_heap_in_lds(L, nnodes);
  const size_t lds = mode_qs_bytes(L) + (in_lds ? mode_heap_bytes(nnodes) : 0);
  hipLaunchKernelGGL(k_mode_search, dim3(nq), dim3(kBlock), lds, st, X, sax, L, Q,
                     paa, kq, nnodes, nd_link, nd_split, nlb, nleaf, leaf_pos, leaf_size, leaf_node, ids, lb, heap_g,
                     in_lds ? 1 : 0, bfs, mp, cap, cand, cand_cnt, bound, keys, nwork, info, counters, mstate, vtop,
                     vcap);
}

}  // namespace hkg
