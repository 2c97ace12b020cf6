// Batched iSAX path (HKG_PATH_BATCHED with a batch of queries): the
// candidate (query, leaf) pairs are ordered by leaf, and one wave evaluates
// the iSAX lower bound (minidist_paa_to_isax_raw_SIMD, sax.c:310-479) of a
// leaf's series for up to 64 of its queries at once -- one query per lane,
// the leaf's iSAX words read once as wave-uniform (scalar) values, the
// breakpoints of their symbols from a table in scalar loads -- so a batch
// reads each candidate leaf's words once per 64 queries instead of once per
// query.  The wave refines its own survivors (bound within the query's k-th
// distance bound after the approximate stage) as they accumulate, one
// (query, series) pair per lane with early abandon against that bound
// (exact_de_knn_read_worker_lf_sims, hercules_query_engine.c:1960-2030);
// series within the bound go to the query's candidate buffer, and k_select
// keeps the k smallest distinct distances.
#include <float.h>
#include <string.h>

#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <vector>

#include "hkg_ed.h"
#include "hkg_engine.h"
#include "hkg_kernels.h"

namespace hkg {

namespace {

constexpr int kStage = 16;               // survivors staged per lane before a flush
constexpr int kIsaxWaves = 4;            // waves per filter block
// The leaf-ordered filter streams each leaf's iSAX words once for all the
// queries that need the leaf, one per lane; with few queries per leaf its
// waves run mostly empty and the per-item filter is faster.  Measured on
// MI355X (series of 256, leaves of 10000, k ~ U{1..100}): it wins from 1024
// queries at 1811 leaves and from 512 at 17861, and loses up to 2048 queries
// at 188 leaves; hence a batch of at least kIsaxMinBatch queries and
// kIsaxMinPairs (query, leaf) pairs takes it.
constexpr uint32_t kIsaxMinBatch = 512;
constexpr uint64_t kIsaxMinPairs = 1536ull * 1024;


// dense list of the candidate pairs keyed by leaf: key = leaf when the pair
// survives the first bound and is not one of the query's best leaves (those
// the approximate stage scanned), else all ones; value = q * nleaf + leaf
__global__ void __launch_bounds__(kBlock) k_collect_by_leaf(uint32_t nq, uint32_t nleaf, const float *__restrict__ lb,
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
    bool is_best = false;
#pragma unroll
    for (int m = 0; m < kBestLeaves; ++m) is_best |= leaf == best[q * kBestLeaves + m];
    const uint32_t bb = bound[q];
    keep = !is_best && (bb >= kInfBits || lb[idx] <= __uint_as_float(bb) * (1.0f + kLbRel) + kLbAbs);
    keys[idx] = keep ? leaf : 0xFFFFFFFFu;
    vals[idx] = idx;
  }
  const uint64_t mask = __ballot(keep);
  if ((threadIdx.x & 63) == 0 && mask) atomicAdd(nwork, (uint32_t)__popcll(mask));
}

// minidist_paa_to_isax_raw_SIMD for one 16-symbol word whose symbols (and so
// breakpoints) are the same for every lane, against the lane's own PAA: the
// same operations in the same order as lb_sax_word (max(lo - p, p - hi, 0)
// squared is bit-identical to the reference's branches, as fl(hi - p) =
// -fl(p - hi))
__device__ __forceinline__ float lb_sax_uniform(const float *p, uint4 w, const float2 *__restrict__ lohi, float ratio)
{
  const uint32_t words[4] = {w.x, w.y, w.z, w.w};
  float t[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const uint32_t v = (words[i >> 2] >> ((i & 3) * 8)) & 0xFFu;
    const float2 lh = lohi[v];
    const float m = fmaxf(fmaxf(lh.x - p[i], p[i] - lh.y), 0.0f);
    t[i] = m * m;
  }
  const float w0 = t[0] + t[8], w1 = t[1] + t[9], w2 = t[2] + t[10], w3 = t[3] + t[11];
  const float w4 = t[4] + t[12], w5 = t[5] + t[13], w6 = t[6] + t[14], w7 = t[7] + t[15];
  const float s = ((w0 + w1) + (w2 + w3)) + ((w4 + w5) + (w6 + w7));
  return s * ratio;
}

// The filter.  lohi: per symbol {lower, upper} breakpoint, with the
// sentinels of minidist_paa_to_isax_raw_SIMD for the first and the last
// symbol (a kernel argument, so that its loads stay scalar with a base held
// in registers).  A wave takes chunks of 64 consecutive entries of the
// leaf-ordered pair list (one pair per lane) from a shared counter; for each
// distinct leaf of its chunk it streams the leaf's iSAX words once, every
// lane of that leaf testing its own query.  Survivors are staged per lane in
// LDS; when a lane's stage fills (and at the end of the chunk) the wave
// compacts all staged (series, query) pairs and refines them 64 at a time,
// one per lane: streamed in 128-byte steps in the reference's order
// (ts.c:83-109), abandoned once the partial sum passes the query's bound
// (partial sums only grow); series within the bound go to the query's
// candidate buffer.  The refinement's row reads overlap other waves' bound
// arithmetic.
__global__ void __launch_bounds__(kIsaxWaves * 64) k_isax_filter(
    const float *__restrict__ X, const uint8_t *__restrict__ sax, const float2 *__restrict__ lohi, uint32_t L,
    const float *__restrict__ Q, const float *__restrict__ paa, uint32_t nleaf, const uint64_t *__restrict__ leaf_pos,
    const uint32_t *__restrict__ leaf_size, const uint32_t *__restrict__ ids, const uint32_t *__restrict__ work,
    const uint32_t *__restrict__ nwork_p, uint32_t *__restrict__ chunk_next, const float *__restrict__ thr,
    uint32_t cap, uint64_t *__restrict__ cand, uint32_t *__restrict__ cand_cnt, const uint32_t *__restrict__ bound,
    unsigned long long *__restrict__ counters)
{
  __shared__ uint32_t stage[kIsaxWaves][kStage][64];     // [wave][slot][lane]: staged series
  __shared__ uint32_t comp_row[kIsaxWaves][kStage * 64];  // the wave's staged pairs, compacted
  __shared__ uint8_t comp_own[kIsaxWaves][kStage * 64];   // lane (query) of each
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t nwork = *nwork_p;
  const uint32_t nchunks = (nwork + 63) / 64;
  const float ratio = (float)(L / kPaaSegments);
  const int nch = (int)(L >> 3);
  unsigned long long checked = 0, words = 0, refined = 0;
  for (;;) {
    uint32_t c = 0;
    if (lane == 0) c = atomicAdd(chunk_next, 1u);
    c = __builtin_amdgcn_readfirstlane(c);
    if (c >= nchunks) break;
    const uint32_t e = c * 64 + lane;
    const bool valid = e < nwork;
    const uint32_t v = valid ? work[e] : 0u;
    const uint32_t q = v / nleaf, leaf = v - q * nleaf;
    float p[kPaaSegments];
#pragma unroll
    for (int i = 0; i < kPaaSegments; i += 4) {
      const float4 f = valid ? *reinterpret_cast<const float4 *>(paa + (uint64_t)q * kPaaSegments + i)
                             : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      p[i] = f.x;
      p[i + 1] = f.y;
      p[i + 2] = f.z;
      p[i + 3] = f.w;
    }
    // nothing passes for an empty lane (bounds are >= 0)
    const float t = valid ? thr[q] : -1.0f;
    int n = 0;
    // refine every staged pair of the wave, then empty the stages
    auto refine_staged = [&]() {
      int off = n;  // inclusive prefix sum of the lanes' counts
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(off, d);
        if (lane >= d) off += o;
      }
      const int total = __shfl(off, 63);
      off -= n;
      for (int j = 0; j < n; ++j) {
        comp_row[wave][off + j] = stage[wave][j][lane];
        comp_own[wave][off + j] = (uint8_t)lane;
      }
      __builtin_amdgcn_wave_barrier();
      for (int base = 0; base < total; base += 64) {
        const int i = base + lane;
        const int own = i < total ? (int)comp_own[wave][i] : lane;
        const uint32_t qq = (uint32_t)__shfl((int)q, own);
        if (i < total) {
          const uint64_t row = comp_row[wave][i];
          const uint32_t bnd = __atomic_load_n(&bound[qq], __ATOMIC_RELAXED);
          const float4 *xp = reinterpret_cast<const float4 *>(X + row * L);
          const float4 *q4 = reinterpret_cast<const float4 *>(Q + (uint64_t)qq * L);
          float acc = 0.0f;
          bool alive = true;
          int ch = 0;
          for (; ch + 4 <= nch && alive; ch += 4) {
            float4 xv[8], qv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              xv[u] = ld_stream(reinterpret_cast<const float *>(xp + 2 * ch + u));
              qv[u] = q4[2 * ch + u];
            }
#pragma unroll
            for (int sub = 0; sub < 4; ++sub)
              acc = acc + ed_chunk(qv[2 * sub], qv[2 * sub + 1], xv[2 * sub], xv[2 * sub + 1]);
            alive = __float_as_uint(acc) <= bnd;
          }
          for (; ch < nch && alive; ++ch) {
            acc = acc + ed_chunk(q4[2 * ch], q4[2 * ch + 1], ld_stream(reinterpret_cast<const float *>(xp + 2 * ch)),
                                 ld_stream(reinterpret_cast<const float *>(xp + 2 * ch + 1)));
            alive = __float_as_uint(acc) <= bnd;
          }
          if (alive) {
            const uint32_t slot = atomicAdd(&cand_cnt[qq], 1u);
            if (slot < cap) cand[(uint64_t)qq * cap + slot] = make_key(acc, ids ? ids[row] : (uint32_t)row);
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      refined += lane == 0 ? (unsigned long long)total : 0ull;
      n = 0;
    };
    uint64_t rem = __ballot(valid);
    while (rem) {
      const int lead = __ffsll((long long)rem) - 1;
      const uint32_t lf = __builtin_amdgcn_readfirstlane(__shfl(leaf, lead));
      const bool mine = valid && leaf == lf;
      rem &= ~__ballot(mine);
      const uint64_t pos = leaf_pos[lf];
      const uint32_t size = leaf_size[lf];
      const uint4 *w4 = reinterpret_cast<const uint4 *>(sax + pos * kPaaSegments);
      for (uint32_t r = 0; r < size; ++r) {
        // the word is the same for every lane: keep it (and so its symbols
        // and their breakpoint loads) scalar
        uint4 w = w4[r];
        w.x = __builtin_amdgcn_readfirstlane(w.x);
        w.y = __builtin_amdgcn_readfirstlane(w.y);
        w.z = __builtin_amdgcn_readfirstlane(w.z);
        w.w = __builtin_amdgcn_readfirstlane(w.w);
        const float lbv = lb_sax_uniform(p, w, lohi, ratio);
        if (mine && lbv <= t) {
          stage[wave][n][lane] = (uint32_t)(pos + r);
          ++n;
        }
        if (__ballot(n == kStage)) refine_staged();
      }
      checked += mine ? size : 0u;
      words += lane == 0 ? size : 0u;
    }
    if (__ballot(n > 0)) refine_staged();
  }
  for (int off = 32; off > 0; off >>= 1) checked += __shfl_down(checked, off);
  if (lane == 0) {
    if (checked) atomicAdd(&counters[kCtrSaxChecked], checked);
    if (words) atomicAdd(&counters[kCtrSaxWords], words);
    if (refined) atomicAdd(&counters[kCtrRefined], refined);
  }
}

__global__ void k_isax_thr(uint32_t nq, const uint32_t *__restrict__ bound, float *__restrict__ thr)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nq) {
    const uint32_t bb = bound[i];
    thr[i] = bb >= kInfBits ? FLT_MAX : __uint_as_float(bb) * (1.0f + kLbRel) + kLbAbs;
  }
}

}  // namespace

hkg_status upload_isax_tables(DevIndex &dv)
{
  float bp[255];
  host_breakpoints(bp);
  float2 lohi[256];
  for (int v = 0; v < 256; ++v)
    lohi[v] = make_float2(v == 0 ? -2000000.0f : bp[v - 1], v == 255 ? 2000000.0f : bp[v]);
  if (!dv.lohi && hipMalloc((void **)&dv.lohi, sizeof(lohi)) != hipSuccess) {
    (void)hipGetLastError();
    dv.lohi = nullptr;
    set_error("no device memory for the iSAX breakpoint table");
    return HKG_ENOMEM;
  }
  HKG_HIP_TRY(hipMemcpy(dv.lohi, lohi, sizeof(lohi), hipMemcpyHostToDevice));
  dv.device_bytes += sizeof(lohi);
  return HKG_OK;
}

bool isax_batched(uint32_t nq, uint32_t nleaf)
{
  return nq >= kIsaxMinBatch && (uint64_t)nleaf * nq >= kIsaxMinPairs;
}

void QueryBuffers::free_isax()
{
  if (is_thr) (void)hipFree(is_thr);
  is_thr = nullptr;
}

hkg_status Index::run_isax_batch(const float *dQ, const uint32_t *dk, uint32_t nq, uint32_t kmax, hipStream_t st,
                                 QueryBuffers **used, Counters *ctr, Prof *pf, const uint32_t *seed_host,
                                 bool use_graph)
{
  const uint32_t L = dev.L, nleaf = dev.nleaf;
  const uint32_t cap = cand_cap(HKG_PATH_BATCHED, std::max(nq, call_settings().min_buffer_batch), kmax);
  QueryBuffers *b = nullptr;
  hkg_status rc = get_buffers(nq, kmax, cap, &b);
  if (rc != HKG_OK) return rc;
  BufLease lease{this, b};  // handed to the caller on success
  if (!b->is_thr && hipMalloc((void **)&b->is_thr, (size_t)b->B * 4) != hipSuccess) {
    (void)hipGetLastError();
    b->is_thr = nullptr;
    set_error("no device memory for the iSAX filter thresholds");
    return HKG_ENOMEM;
  }
  HKG_HIP_TRY(hipMemcpyAsync(b->Q, dQ, (size_t)nq * L * 4, hipMemcpyDeviceToDevice, st));
  HKG_HIP_TRY(hipMemcpyAsync(b->kq, dk, (size_t)nq * 4, hipMemcpyDeviceToDevice, st));
  if (seed_host) HKG_HIP_TRY(hipMemcpyAsync(b->seed, seed_host, (size_t)nq * 4, hipMemcpyHostToDevice, st));
  uint32_t *nwork = b->misc + 0, *chunk_next = b->misc + 1, *overflow = b->misc + 2;
  int d = 0, cus = 0;
  (void)hipGetDevice(&d);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d);
  const uint32_t filter_blocks = (uint32_t)std::max(cus, 1) * 4;
  auto enqueue = [&](Prof *p) -> hkg_status {
    launch_reset(st, nq, b->bound, b->cand_cnt, seed_host ? b->seed : nullptr);
    launch_fill(st, b->misc, 8, 0);
    launch_check_k(st, nq, kmax, b->kq, b->misc + 3);
    launch_fill(st, reinterpret_cast<uint32_t *>(b->counters), 2 * kNumCounters, 0);
    if (p) p->mark(st, -1);
    launch_prep(st, b->Q, nq, L, b->paa, b->psum, b->qn2);
    if (p) p->mark(st, HKG_STAGE_PREP);
    launch_eapca(st, nq, nleaf, L, dev.leaf_seg_off, dev.seg_end, dev.seg_ind, b->psum, b->lb);
    if (p) p->mark(st, HKG_STAGE_EAPCA);
    launch_best_leaf(st, nq, nleaf, b->lb, b->best);
    if (p) p->mark(st, HKG_STAGE_SELECT_LEAF);
    launch_scan_best(st, nq, dev.X, L, b->Q, b->kq, b->best, dev.leaf_pos, dev.leaf_size, dev.ids, b->cap, b->cand,
                     b->cand_cnt, b->bound, b->counters);
    if (p) p->mark(st, HKG_STAGE_APPROX);
    const uint32_t pairs = nq * nleaf;
    uint32_t *keys2 = b->work_keys + b->work_cap, *vals2 = b->work_vals + b->work_cap;
    hipLaunchKernelGGL(k_collect_by_leaf, dim3((nleaf + kBlock - 1) / kBlock, nq), dim3(kBlock), 0, st, nq, nleaf,
                       b->lb, b->best, b->bound, b->work_keys, b->work_vals, nwork);
    HKG_HIP_TRY(launch_sort_work(st, b->sort_tmp, b->sort_bytes, b->work_keys, keys2, b->work_vals, vals2, pairs));
    hipLaunchKernelGGL(k_isax_thr, dim3((nq + 255) / 256), dim3(256), 0, st, nq, b->bound, b->is_thr);
    if (p) p->mark(st, HKG_STAGE_COLLECT);
    hipLaunchKernelGGL(k_isax_filter, dim3(filter_blocks), dim3(kIsaxWaves * 64), 0, st, dev.X, dev.sax, dev.lohi, L,
                       b->Q, b->paa, nleaf, dev.leaf_pos, dev.leaf_size, dev.ids, vals2, nwork, chunk_next, b->is_thr,
                       b->cap, b->cand, b->cand_cnt, b->bound, b->counters);
    if (p) p->mark(st, HKG_STAGE_REFINE);
    launch_select(st, nq, b->kmax, b->kq, b->cap, b->cand, b->cand_cnt, idmap64, b->out_ids, b->out_d2, overflow);
    if (p) p->mark(st, HKG_STAGE_TOPK);
    return HKG_OK;
  };
  if (use_graph && !seed_host)
    rc = run_captured(*b, 3, nq, 0, st, pf, enqueue);
  else
    rc = enqueue(pf);
  if (rc != HKG_OK) return rc;
  HKG_HIP_TRY(hipGetLastError());
  if (ctr) {
    HKG_HIP_TRY(hipMemcpyAsync(ctr->v, b->counters, sizeof(ctr->v), hipMemcpyDeviceToHost, st));
    HKG_HIP_TRY(hipMemcpyAsync(&ctr->overflow, overflow, 4, hipMemcpyDeviceToHost, st));
    HKG_HIP_TRY(hipMemcpyAsync(&ctr->nwork, nwork, 4, hipMemcpyDeviceToHost, st));
    HKG_HIP_TRY(hipMemcpyAsync(&ctr->bad_k, b->misc + 3, 4, hipMemcpyDeviceToHost, st));
  }
  *used = b;
  lease.b = nullptr;
  return HKG_OK;
}

}  // namespace hkg
