// Engine objects behind the C ABI handles.
#pragma once

#include <atomic>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "hkg_internal.h"
#include "hkg_kernels.h"

namespace hkg {

const std::string &last_error();

struct Counters {
  unsigned long long v[kNumCounters] = {};
  uint32_t overflow = 0;
  uint32_t nwork = 0;
  uint32_t bad_k = 0;        // per-query k outside [1, kmax] (clamped on the device)
  uint64_t bf16_rows = 0;    // rows the BF16 collecting pass read (summed over block groups)
  uint64_t bf16_rows_bound = 0;  // rows the BF16 histogram pass read
  bool isax = false;         // the batch ran the leaf-ordered iSAX filter
};

// per-stage timing of one batch (profile mode): a one-wave kernel stamps the
// device's constant-rate wall clock into d_ts at every stage boundary, on the
// batch's stream, so consecutive stamps bracket one stage's kernels.  The
// stamps are ordinary kernels, so a captured graph records them as well and
// its replays time their stages without eager launches.
struct Prof {
  static constexpr int kMarks = 12;
  unsigned long long *d_ts = nullptr;  // device: one stamp per mark
  int stage[kMarks];
  int n = 0;
  bool ok = false;
  double ticks_per_ms = 0.0;
  // a captured graph's own Prof, whose stamps its replay writes, read in
  // place of this one's
  const Prof *src = nullptr;
  bool init()
  {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0 ||
        hipMalloc((void **)&d_ts, kMarks * sizeof(unsigned long long)) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    ticks_per_ms = (double)khz;
    ok = true;
    return true;
  }
  void mark(hipStream_t st, int s)
  {
    if (n < kMarks) {
      launch_stamp(st, d_ts + n);
      stage[n++] = s;
    }
  }
  // after the batch's stream has finished
  void accumulate(double *ms, uint64_t *launches) const
  {
    const Prof &p = src ? *src : *this;
    if (p.n < 2) return;
    unsigned long long ts[kMarks];
    if (hipMemcpy(ts, p.d_ts, p.n * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess) return;
    for (int i = 1; i < p.n; ++i) {
      if (p.stage[i] < 0 || ts[i] < ts[i - 1]) continue;
      ms[p.stage[i]] += (double)(ts[i] - ts[i - 1]) / p.ticks_per_ms;
      launches[p.stage[i]] += 1;
    }
  }
  ~Prof()
  {
    if (d_ts) (void)hipFree(d_ts);
  }
};

// Device scratch of one batch shape.  All per-query arrays are B long.
struct QueryBuffers {
  uint32_t B = 0, kmax = 0, cap = 0, work_cap = 0;
  uint64_t bytes = 0;
  float *Q = nullptr;           // B x L
  uint32_t *kq = nullptr;       // B
  float *paa = nullptr;         // B x 16
  double *psum = nullptr;       // B x 2(L+1)
  double *qn2 = nullptr;        // B
  float *lb = nullptr;          // B x nleaf
  uint32_t *best = nullptr;     // B
  uint32_t *bound = nullptr;    // B, FP32 bits of the current k-th distance bound
  uint64_t *cand = nullptr;     // B x cap candidate keys
  uint32_t *cand_cnt = nullptr; // B
  uint32_t *work_keys = nullptr; // 2 x B x nleaf: EAPCA bound bits (sort keys), double buffer
  uint32_t *work_vals = nullptr; // 2 x B x nleaf: pair index q * nleaf + leaf
  void *sort_tmp = nullptr;
  size_t sort_bytes = 0;
  uint32_t *seed = nullptr;      // B: k-th distance bound seeds (overflow re-runs)
  uint32_t *misc = nullptr;     // [0] nwork [1] work_next [2] overflow
  unsigned long long *counters = nullptr;
  uint64_t *out_ids = nullptr;  // B x kmax
  float *out_d2 = nullptr;      // B x kmax
  // BF16 path scratch
  uint32_t *bf_cnt = nullptr;   // B survivor counters
  uint64_t *bf_surv = nullptr;  // B x bf_cap survivors: (BF16 bound bits << 32) | row
  uint32_t bf_cap = 0;
  uint16_t *Qb = nullptr;       // B x L bf16 bits
  float4 *qst = nullptr;        // B x {|q|^2, |q - q~|, |q~|, 0}
  uint32_t *qperm = nullptr;    // 4 x B: batch order by first bound, then sort scratch
  void *bf_sort_tmp = nullptr;
  size_t bf_sort_bytes = 0;
  uint8_t *need = nullptr;      // groups x nleaf: leaf needed by a block group
  uint32_t *tiles = nullptr;    // groups x ntiles: row tiles of each block group
  uint32_t *ntiles_g = nullptr; // groups
  uint32_t *thr = nullptr;      // B: filter threshold (float bits)
  uint32_t *hist = nullptr;     // B x kBf16Bins: counts of series by BF16 upper bound
  uint8_t *slot_need8 = nullptr;  // wave slots x nleaf: leaf needed by a filter wave's queries
  uint32_t *slot_need = nullptr;  // ntiles x words: per row tile, one bit per wave slot that needs it
  uint4 *bf_region = nullptr;   // per filter block: survivor records (header + 16 values xn2 - 2 s)
  uint32_t *bf_region_cnt = nullptr;  // per filter wave: pre-test passes in its region slice
  uint32_t *bf_prog = nullptr;  // per filter block: tiles done (sibling blocks keep pace)
  void free_bf16();
  // batched iSAX path scratch
  float *is_thr = nullptr;      // B: filter threshold
  void free_isax();
  // approximate modes
  float *nlb = nullptr;         // B x nnodes exact node bounds
  uint64_t *heap = nullptr;     // B x (nnodes + 1) traversal heaps (when too big for LDS)
  uint32_t *bfs = nullptr;      // B x 2 (nnodes + 1) level lists of the candidate walk
  uint32_t *info = nullptr;     // B x 4 per-query traversal counts
  uint32_t *mstate = nullptr;   // B x kModeState: queries k_mode_search hands to k_mode_count / k_mode_finish
  uint64_t *vtop = nullptr;     // B x kmax: their answers after the approximate phase
  unsigned long long *lab = nullptr;  // B: their iSAX counts
  bool busy = false;            // held by a query call (guarded by Index::mu_bufs)
  uint64_t last_use = 0;        // Index::buf_clock at the last lease (guarded by Index::mu_bufs)
  // captured stage sequence of a batched path (FP32 or BF16), and the shape
  // it was captured for
  hipGraphExec_t graph_exec = nullptr;
  hipStream_t graph_stream = nullptr;
  uint32_t graph_kind = 0;      // 1: FP32 batched, 2: BF16
  uint32_t graph_nq = 0;
  uint32_t graph_flags = 0;     // path-specific (BF16: the bound pass)
  uint32_t graph_fgrid = 0;     // filter + refine grid the graph was captured with
  bool graph_profiled = false;  // stage events recorded by the graph
  std::unique_ptr<Prof> graph_prof;
  ~QueryBuffers();
};

struct Index {
  DevIndex dev;
  hipStream_t stream = nullptr;
  uint64_t *idmap64 = nullptr;  // rank -> public id (when an id map is attached)
  // Buffer sets, keyed by shape.  A query call holds each set it uses
  // (busy) until its results are copied out, so concurrent calls on other
  // streams and nested re-runs never share one.
  std::mutex mu_bufs, mu_bf16;
  bool bf16_ready = false;      // BF16 copy, norms and tile tables built (guarded by mu_bf16)
  // key: (queries, kmax, candidate entries per query)
  std::multimap<std::tuple<uint32_t, uint32_t, uint32_t>, std::unique_ptr<QueryBuffers>> bufs;
  uint64_t buf_clock = 0;  // lease counter for the least-recently-used eviction (guarded by mu_bufs)
  uint64_t sets_allocated = 0, sets_evicted = 0;  // buffer sets made and dropped (guarded by mu_bufs)
  // launch shapes chosen by autotune (0: the defaults): queries per batch
  // of each path, and k_filter_refine blocks per CU
  std::atomic<uint32_t> tuned_batch[HKG_PATH_BF16 + 1] = {};
  std::atomic<uint32_t> tuned_fblocks{0};
  std::atomic<uint32_t> tuned_sblocks{0};  // full-scan blocks per CU
  std::atomic<uint32_t> tuned_filter{0};   // batched path's filter at its tuned batch (1 items, 2 leaf; 0 none)
  std::atomic<uint32_t> tuned_qt{0};       // BF16 filter query tiles per wave (0: kBf16QueryTiles)
  uint32_t bf16_qt() const
  {
    const uint32_t t = tuned_qt.load(std::memory_order_relaxed);
    return t ? t : kBf16QueryTiles;
  }

  ~Index();
  hkg_status init_device(int device);
  hkg_status upload_tree();
  hkg_status set_ids(const uint64_t *ids, uint64_t n);
  hkg_status make_bf16();
  hkg_status alloc_buffers(uint32_t B, uint32_t kmax, uint32_t cap, QueryBuffers **out);
  // the BF16 scratch of a buffer set, allocated on first use
  hkg_status ensure_bf16_scratch(QueryBuffers &b);
  // allocate `count` buffer sets of the shape a batch of B queries at kmax
  // leases on `path` (a server's workers, before their first batch)
  hkg_status prepare_buffers(hkg_path path, uint32_t B, uint32_t kmax, uint32_t count);
  hkg_status get_buffers(uint32_t B, uint32_t kmax, uint32_t cap, QueryBuffers **out);
  void release_buffers(QueryBuffers *b);
  // device scratch one query of a batch needs on a path (buffer sets)
  uint64_t scratch_per_query(hkg_path path, uint32_t kmax) const;
  uint32_t cand_cap(hkg_path path, uint32_t nq, uint32_t kmax) const;
  void enqueue_index_batch(QueryBuffers &b, uint32_t nq, hipStream_t st, Prof *pf, bool seeded);
  void enqueue_scan_batch(QueryBuffers &b, uint32_t nq, hipStream_t st, Prof *pf);
  hkg_status enqueue_mode_batch(QueryBuffers &b, uint32_t nq, hipStream_t st, Prof *pf, bool seeded);
  hkg_status run_batch(hkg_path path, const float *dQ, const uint32_t *dk, uint32_t nq, uint32_t kmax,
                       bool use_graph, hipStream_t st, QueryBuffers **used, Counters *ctr, Prof *pf,
                       const uint32_t *seed_host = nullptr);
  hkg_status run_bf16_batch(const float *dQ, const uint32_t *dk, uint32_t nq, uint32_t kmax, hipStream_t st,
                            QueryBuffers **used, Counters *ctr, Prof *pf, const uint32_t *seed_host,
                            bool bound_pass, bool use_graph = false);
  // the batched FP32 path: leaf-ordered iSAX filter + refinement in bound
  // order (hkg_isax.hip)
  hkg_status run_isax_batch(const float *dQ, const uint32_t *dk, uint32_t nq, uint32_t kmax, hipStream_t st,
                            QueryBuffers **used, Counters *ctr, Prof *pf, const uint32_t *seed_host,
                            bool use_graph = false);
  // launch the stage sequence `enqueue` through the buffer set's captured
  // graph (captured again when kind, nq, flags, grid, stream or profiling
  // change); with pf, its stage events are the graph's own
  hkg_status run_captured(QueryBuffers &b, uint32_t kind, uint32_t nq, uint32_t flags, hipStream_t st, Prof *pf,
                          const std::function<hkg_status(Prof *)> &enqueue);
  uint64_t total_scan_bytes(uint32_t nq, uint32_t batch) const;
  uint64_t total_segs() const;
  // persistent grid of k_filter_refine for the running call
  uint32_t filter_grid() const;
  // grid.x of the full scan for the running call (per query group)
  uint32_t scan_grid(uint32_t nq) const;
  // whether an exact batch of nq on the batched path takes the leaf-ordered
  // iSAX filter (the call's batched_filter, else by batch and index size)
  bool leaf_ordered(uint32_t nq) const;
  hkg_status autotune(const float *Q, uint32_t nq, const uint32_t *k, uint32_t kmax, uint32_t path,
                      hkg_tuning *out);
  hkg_status fix_overflow(QueryBuffers &b, uint32_t nq, uint32_t kmax, hipStream_t st, uint32_t *retries);
  hkg_status fix_bf16_overflow(QueryBuffers &b, uint32_t nq, uint32_t kmax, hipStream_t st, uint32_t *retries);
  hkg_status knn_device(const float *dQ, uint32_t nq, const uint32_t *dk, uint32_t kmax, const hkg_query_opts *qo,
                        uint64_t *dids, float *dd2, hipStream_t st, hkg_query_stats *stats);
  hkg_status knn_host(const float *Q, uint32_t nq, const uint32_t *k, uint32_t kmax, const hkg_query_opts *qo,
                      uint64_t *ids, float *d2, hkg_query_stats *stats);
};

// Settings of the query call running on this thread (from hkg_query_opts);
// nested calls save and restore them.
struct CallSettings {
  uint32_t cap_override = 0;  // candidate (or survivor) entries per query requested by the caller
  uint32_t bound_pass = 0;    // hkg_query_opts.bf16_bound_pass
  uint32_t min_buffer_batch = 0;  // size buffer sets for at least this many queries (server workers)
  bool mode = false;          // an approximate mode (epsilon > 0 or nprobes > 0) is on
  uint32_t fblocks = 0;       // k_filter_refine blocks per CU (0: the index's tuned or default)
  uint32_t sblocks = 0;       // full-scan blocks per CU (0: the index's tuned or default)
  uint32_t batched_filter = 0;  // hkg_query_opts.batched_filter
  uint32_t bf16_qt = 0;       // BF16 filter query tiles per wave of this call (0: the default)
  ModeParams mp{};            // its parameters (nleaves / nseries filled per index)
};
CallSettings &call_settings();
// BF16 path scratch of one query beyond the FP32 buffer set
uint64_t bf16_scratch_per_query(uint32_t L, uint64_t nleaf, uint64_t ntiles);
// queries of one BF16 filter block group (a batch's unit of work) with qt
// query tiles per wave, and the tile counts compiled for a length
uint32_t bf16_group_queries(uint32_t L, uint32_t qt);
bool bf16_query_tiles_ok(uint32_t L, uint32_t qt);
// batched iSAX path (hkg_isax.hip): whether a batch of nq queries over nleaf
// leaves takes the leaf-ordered filter by default, and its breakpoint table
// (once per device)
bool isax_batched(uint32_t nq, uint32_t nleaf);
hkg_status upload_isax_tables(DevIndex &dv);

// releases a held buffer set when it goes out of scope
struct BufLease {
  Index *ix;
  QueryBuffers *b;
  ~BufLease()
  {
    if (b) ix->release_buffers(b);
  }
};

// device address of the no-allocation entry point's scratch array, looked
// up once per device (hkg_noalloc.hip)
hkg_status noalloc_setup(int device);

hkg_status open_dir(const char *dir, const hkg_options *opt, Index **out);
hkg_status open_raw(const float *X, uint64_t n, uint32_t L, const hkg_options *opt, Index **out);
// full-scan grid.x for nq queries of length L at blocks_per_cu resident
// blocks per CU
uint32_t scan_grid_x(uint32_t nq, uint32_t L, uint32_t blocks_per_cu);
hkg_status build_device(float *dX, uint64_t n, uint32_t L, const hkg_build_opts &bo, Index **out);
hkg_status write_dir(const Index &ix, const char *dir);

}  // namespace hkg
