// Launchers of the query kernels (hkg_kernels.hip).
#pragma once

#include "hkg_internal.h"

namespace hkg {

constexpr int kRefineTile = 2048;        // floats of chunk sums per wave (refine)
constexpr int kRefineTileFilter = 1024;  // the same in k_filter_refine (four blocks per CU)
constexpr uint32_t kFilterBlocksPerCU = 4;  // persistent grid of k_filter_refine
constexpr int kScanQ = 4;                // queries per block in the full scan
constexpr uint32_t kScanBlocksPerCU = 3;  // default persistent grid of the full scan
// queries per pass of the full scan for a batch of nq series of length L:
// up to kScanQ, fewer where their rows would not fit 64 KB of LDS
inline uint32_t scan_group_size(uint32_t nq, uint32_t L)
{
  const uint32_t fit = L <= 4096 ? 4u : (L <= 8192 ? 2u : 1u);
  const uint32_t g = nq <= 2 ? (nq ? nq : 1) : (uint32_t)kScanQ;
  return g < fit ? g : fit;
}
constexpr size_t kEapcaLdsBytes = 48 * 1024;
constexpr int kWorkChunk = 8;            // work items taken per atomic
constexpr int kFilterRows = 1024;        // rows per iSAX filter step of k_filter_refine
constexpr int kBestLeaves = 6;           // leaves per query scanned in the approximate stage
constexpr int kBf16Bins = 32;            // threshold histogram bins per query (BF16 path)
constexpr uint32_t kBf16QueryTiles = 2;   // default 32-query MFMA tiles per BF16 filter wave

// device counters (unsigned long long each)
enum Counter : int {
  kCtrLeaves = 0,      // (query, leaf) pairs after the EAPCA bound (excl. best leaf)
  kCtrSaxChecked = 1,  // series whose iSAX bound was evaluated
  kCtrRefined = 2,     // series whose exact distance was computed
  kCtrBf16 = 3,        // series whose BF16 bound was evaluated
  kCtrApprox = 4,      // series of the best leaves scanned in the approximate stage
  kCtrModeFirst = 5,   // approximate modes: leaves scanned before the candidate stage
  kCtrModeCand = 6,    // approximate modes: candidate leaves
  kCtrModeScanned = 7, // approximate modes: candidate leaves scanned or refined
  kCtrModeSeq = 8,     // approximate modes: queries whose candidates went sequential
  kCtrApproxLeaves = 9,  // (query, leaf) pairs scanned in the approximate stage
  kCtrLeafPairs = 10,  // BF16 path: (query, leaf) pairs within the k-th bound after the approximate stage
  kCtrSaxWords = 11,   // batched iSAX path: iSAX words streamed (once per leaf per wave chunk)
  kCtrBf16Skip = 12,   // BF16 path: (query, series) pairs of tiles a filter wave skipped (no query needs them)
  kCtrApproxRanked = 13,  // BF16 path: series of the approximate stage's leaves ranked (within the iSAX bound)
  kNumCounters = 14
};

// parameters of the approximate modes (hkg_query_opts)
struct ModeParams {
  float epsilon, eapca_th, sax_th;
  uint32_t budget, nprobes;
  uint32_t nleaves;   // leaves with rows (the reference's leaves_counter)
  uint64_t nseries;   // rows under the root
};
// the traversal heap lives in LDS when it and the query fit in this many
// bytes (one block per CU), else in per-query scratch
constexpr size_t kModeLdsMax = 144 * 1024;
bool mode_heap_in_lds(uint32_t L, uint32_t nnodes);

hipError_t upload_breakpoints_dev();
// the query kernels that hold a query (or a group of them) in LDS may use
// up to kQueryLdsMax bytes of dynamic LDS (series up to kMaxLength)
constexpr size_t kQueryLdsMax = 96 * 1024;
void set_lds_limits();
void set_mode_lds_limits();

void launch_reset(hipStream_t st, uint32_t nq, uint32_t *bound, uint32_t *cand_cnt, const uint32_t *seed);
// per-query k in place: values outside [1, kmax] are clamped (so no kernel
// sees them) and counted in *bad, which the host turns into HKG_EINVAL
void launch_check_k(hipStream_t st, uint32_t nq, uint32_t kmax, uint32_t *kq, uint32_t *bad);
void launch_fill(hipStream_t st, uint32_t *p, uint32_t n, uint32_t v);
// one wave writes wall_clock64() (hipDeviceAttributeWallClockRate ticks) to *t
void launch_stamp(hipStream_t st, unsigned long long *t);
void launch_prep(hipStream_t st, const float *Q, uint32_t nq, uint32_t L, float *paa, double *psum,
                 double *qn2);
void launch_eapca(hipStream_t st, uint32_t nq, uint32_t nleaf, uint32_t L, const uint32_t *seg_off,
                  const int16_t *seg_end, const float4 *ind, const double *psum, float *lb);
void launch_best_leaf(hipStream_t st, uint32_t nq, uint32_t nleaf, const float *lb, uint32_t *best);
void launch_scan_best(hipStream_t st, uint32_t nq, const float *X, uint32_t L, const float *Q,
                      const uint32_t *kq, const uint32_t *best, const uint64_t *leaf_pos,
                      const uint32_t *leaf_size, const uint32_t *ids, uint32_t cap, uint64_t *cand,
                      uint32_t *cand_cnt, uint32_t *bound, unsigned long long *counters);
// BF16 path: the best leaves ranked from the BF16 copy, exact distances only
// for the best-ranked rows (falls back to launch_scan_best when the largest
// leaf's ranking keys exceed kApproxMaxDynLds)
constexpr size_t kApproxMaxDynLds = 64 * 1024;
// the best leaves' rows ranked by a BF16 dot, the best k + 32 of each refined
// exactly: the first leaf of every query, then the others with the rows whose
// iSAX bound (sax words, query PAA) is within the k-th distance so far
void launch_scan_best_bf16(hipStream_t st, uint32_t nq, const float *X, const uint16_t *Xb, const float *xn2,
                           uint32_t max_leaf, uint32_t L, const float *Q, const uint32_t *kq, const uint8_t *sax,
                           const float *paa, const uint32_t *best,
                           const uint64_t *leaf_pos, const uint32_t *leaf_size, const uint32_t *ids, uint32_t cap,
                           uint64_t *cand, uint32_t *cand_cnt, uint32_t *bound, unsigned long long *counters);
void launch_collect(hipStream_t st, uint32_t nq, uint32_t nleaf, const float *lb, const uint32_t *best,
                    const uint32_t *bound, uint32_t *keys, uint32_t *vals, uint32_t *nwork);
size_t sort_work_temp_bytes(uint32_t n);
hipError_t launch_sort_work(hipStream_t st, void *tmp, size_t tmp_bytes, const uint32_t *keys, uint32_t *keys_out,
                            const uint32_t *vals, uint32_t *vals_out, uint32_t n);
void launch_filter_refine(hipStream_t st, uint32_t grid, const float *X, const uint8_t *sax, uint32_t L,
                          const float *Q, const float *paa, const uint32_t *kq, const uint64_t *leaf_pos,
                          const uint32_t *leaf_size, const uint32_t *ids, uint32_t nleaf, const float *lb,
                          const uint32_t *work, const uint32_t *nwork, uint32_t *work_next, uint32_t cap,
                          uint64_t *cand, uint32_t *cand_cnt, uint32_t *bound, unsigned long long *counters);
// approximate modes
void launch_node_lb(hipStream_t st, uint32_t nq, uint32_t nnodes, uint32_t L, const float *Q, const uint2 *nd_seg,
                    const int16_t *seg_end, const float4 *ind, float *nlb);
void launch_mode_keys(hipStream_t st, uint32_t pairs, uint32_t *keys, uint32_t *vals);
void launch_mode_search(hipStream_t st, uint32_t nq, const float *X, const uint8_t *sax, uint32_t L, const float *Q,
                        const float *paa, const uint32_t *kq, uint32_t nnodes, const int4 *nd_link,
                        const int4 *nd_split, const float *nlb, uint32_t nleaf, const uint64_t *leaf_pos,
                        const uint32_t *leaf_size, const int32_t *leaf_node, const uint32_t *ids, const float *lb,
                        uint64_t *heap_g, uint32_t *bfs, const ModeParams &mp, uint32_t cap, uint64_t *cand,
                        uint32_t *cand_cnt, uint32_t *bound, uint32_t *keys, uint32_t *nwork, uint32_t *info,
                        unsigned long long *counters, uint32_t *mstate, uint64_t *vtop, uint32_t vcap);
// the epsilon search's queries whose candidate branch needs the iSAX count
// or is sequential: k_mode_search hands them over (kModeState words per
// query, and its answer so far), k_mode_count counts, k_mode_finish ends them
constexpr int kModeState = 8;
void launch_mode_count(hipStream_t st, uint32_t nq, const uint8_t *sax, uint32_t L, const float *paa, uint32_t nnodes,
                       const float *nlb, uint32_t nleaf, const uint64_t *leaf_pos, const uint32_t *leaf_size,
                       const int32_t *leaf_node, const uint32_t *keys, const uint32_t *mstate, unsigned long long *lab);
void launch_mode_finish(hipStream_t st, uint32_t nq, const float *X, uint32_t L, const float *Q, const uint32_t *kq,
                        uint32_t nnodes, const float *nlb, uint32_t nleaf, const uint64_t *leaf_pos,
                        const uint32_t *leaf_size, const int32_t *leaf_node, const uint32_t *ids, const ModeParams &mp,
                        uint32_t cap, uint64_t *cand, uint32_t *cand_cnt, uint32_t *bound, uint32_t *keys,
                        uint32_t *nwork, uint32_t *info, unsigned long long *counters, const uint32_t *mstate,
                        const uint64_t *vtop, uint32_t vcap, const unsigned long long *lab);
void launch_scan(hipStream_t st, uint32_t grid_x, const float *X, uint64_t n, uint32_t L, const float *Q,
                 uint32_t nq, const uint32_t *kq, const uint32_t *ids, uint32_t cap, uint64_t *cand,
                 uint32_t *cand_cnt, uint32_t *bound);
void launch_bf16_refine(hipStream_t st, uint32_t nq, const float *X, uint32_t L, const float *Q, const uint32_t *kq,
                        uint32_t kmax, const uint32_t *ids, uint32_t cap, const uint64_t *cand,
                        const uint32_t *cand_cnt, uint32_t scap, const uint64_t *surv, uint32_t *scnt,
                        const float *thr0, const uint32_t *bound, const uint64_t *idmap64, uint64_t *out_ids,
                        float *out_d2, uint32_t *overflow, unsigned long long *counters);
void launch_query_bf16(hipStream_t st, uint32_t nq, uint32_t L, const float *Q, uint16_t *Qb, float4 *qst);
void launch_select(hipStream_t st, uint32_t nq, uint32_t kmax, const uint32_t *kq, uint32_t cap,
                   const uint64_t *cand, const uint32_t *cand_cnt, const uint64_t *idmap64, uint64_t *out_ids,
                   float *out_d2, uint32_t *overflow);

}  // namespace hkg
