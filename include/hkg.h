/*
 * hkg.h -- C ABI of the MI355X exact k-NN query engine for Hercules indexes.
 *
 * Plain C types only (no HIP or torch types).  Every entry point returns an
 * hkg_status; none calls exit().  A handle is safe to use from one thread at
 * a time; distinct handles are independent.
 *
 * Reference interfaces replaced (file:line in karimaechihabi/hercules):
 *   hkg_open_dir      hercules_index_read + cache_sax_file
 *                     (code/hercules/src/hercules_index.c:1793-1905, 2342-2351)
 *   hkg_knn           the exact query functions exact_de_knn_search,
 *                     exact_de_knn_search_psq_lf and
 *                     exact_de_knn_search_psq_lf_sims_adaptive
 *                     (code/hercules/include/hercules_query_engine.h:203-207,
 *                     256-261, 284-289), batched over queries and returning
 *                     the results instead of printing them; the drop-in
 *                     definitions of all three (and of exact_ng_knn_search)
 *                     are hercules_amd/integration/hkg_shim.c
 *   hkg_open_raw +
 *   hkg_knn(HKG_PATH_SCAN)  PSCAN's parallel_scan / parallel_scan_process_query
 *                     (code/pscan/src/parallel_scan_utils.c:23-347)
 *   hkg_build         hercules_index_build + hercules_index_write
 *                     (code/hercules/src/hercules_index.c:2354-2929,
 *                     1694-1791): a GPU bulk loader producing the same
 *                     on-disk format (root.idx, leaves_raw.idx,
 *                     leaves_sims.idx)
 *   hkg_autotune      no counterpart: the reference's query-time shape is
 *                     its thread counts, set by hand on the command line
 *   hkg_server_*      no single counterpart: the query file loop
 *                     hercules_knn_query_binary_file
 *                     (code/hercules/src/hercules_file_loaders.c:99-294) as a
 *                     continuous-batching request loop
 *
 * Result semantics follow queue_bounded_sorted_insert
 * (code/hercules/src/hercules_query_engine.c:4034-4082) and its strict-<
 * call sites: per query the k smallest DISTINCT FP32 squared distances in
 * ascending order; among series with equal distance the smallest id is
 * returned.  Ranks that cannot be filled hold d2 = FLT_MAX and
 * id = UINT64_MAX.  Squared distances are bit-identical to the reference's
 * ts_euclidean_distance_SIMD (code/hercules/src/ts.c:83-109).
 */
#ifndef HKG_H
#define HKG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hkg_status {
  HKG_OK = 0,
  HKG_EINVAL = 1,   /* bad argument */
  HKG_ENOMEM = 2,   /* host or device allocation failed */
  HKG_EIO = 3,      /* file missing, unreadable or short */
  HKG_EFORMAT = 4,  /* index file malformed or truncated */
  HKG_EDEVICE = 5,  /* HIP runtime error */
  HKG_EBUSY = 6,    /* serving loop full (admission refused) */
  HKG_ENOTREADY = 7, /* serving ticket not finished yet */
  HKG_EINTERNAL = 8  /* an engine self-check failed */
} hkg_status;

typedef struct hkg_index hkg_index;
typedef struct hkg_server hkg_server;

typedef enum hkg_path {
  HKG_PATH_AUTO = 0,    /* batched index path (BF16 filter where it pays)   */
  HKG_PATH_INDEX = 1,   /* plain FP32 index path, one query per launch set  */
  HKG_PATH_BATCHED = 2, /* batched FP32 index path: persistent (query, leaf)
                           work queue, stages captured in a hipGraph        */
  HKG_PATH_SCAN = 3,    /* PSCAN: exact full scan, no lower bounds          */
  HKG_PATH_BF16 = 4     /* batched index path with the BF16 MFMA certified
                           lower bound in place of the iSAX filter          */
} hkg_path;

typedef struct hkg_options {
  int device;          /* HIP device ordinal                                  */
  int keep_bf16;       /* also keep a BF16 copy of the series (for HKG_PATH_BF16) */
  int verbose;         /* print load statistics to stderr                     */
} hkg_options;

typedef struct hkg_query_opts {
  hkg_path path;
  uint32_t batch;      /* queries per batch (0 = library default)             */
  uint32_t profile;    /* 1: time every stage with events (no graph replay)   */
  uint32_t cand_cap;   /* candidate buffer entries per query (0 = default);
                          when a query overflows it, it is answered again    */
  uint32_t bf16_bound_pass; /* BF16 path: 0 = only when re-running queries
                               whose survivor list overflowed, 1 = always,
                               2 = never (seeded re-runs only)              */
  /* Approximate modes of the reference's query answering (hercules.c:48-51,
     :70-72).  Exact search when epsilon == 0 and nprobes == 0; otherwise the
     answers are the reference's for the same index and settings, and
     `path` is ignored (the modes run their own stage sequence).  Squared
     distances throughout, as in the reference. */
  float epsilon;           /* > 0: epsilon search: leaves are pruned against
                              kth / (1 + epsilon)                            */
  float delta;             /* accepted for the reference's interface; it has
                              no effect there either (r_delta = FLT_MAX,
                              hercules.c:479)                                */
  uint32_t nprobes;        /* > 0: no-guarantee search that scans nprobes
                              leaves (exact_ng_knn_search); epsilon ignored  */
  uint32_t approx_leaves;  /* epsilon search: leaf budget of its approximate
                              phase (--approx-stop-condition; default 80;
                              the first leaf is always scanned, so 0 acts
                              as 1, as in the reference)                     */
  float eapca_threshold;   /* epsilon search: candidate-branch thresholds    */
  float sax_threshold;     /* (--eapca-threshold 0.25, --sax-threshold 0.9)  */
  uint32_t batched_filter; /* HKG_PATH_BATCHED, exact search: the filter of a
                              batch.  0 = the autotuner's choice at its tuned
                              batch, else by batch and index size (the
                              leaf-ordered filter from 512 queries and
                              1.5 M (query, leaf) pairs per batch),
                              1 = per (query, leaf) work items,
                              2 = leaf-ordered: the queries of each leaf
                              together, one query per lane.  Same answers.  */
  uint32_t bf16_query_tiles; /* HKG_PATH_BF16: 32-query MFMA tiles per filter
                              wave (2, 3; 4 for L <= 128), so 128 x this many
                              queries per filter block.  0 = the autotuner's
                              choice, else 2.  Same answers.               */
} hkg_query_opts;

/* stages reported in hkg_query_stats.stage_ms / stage_bytes */
enum {
  HKG_STAGE_PREP = 0,     /* PAA and prefix sums of the queries              */
  HKG_STAGE_EAPCA = 1,    /* EAPCA bound of every (query, leaf)              */
  HKG_STAGE_SELECT_LEAF = 2, /* best leaf per query                          */
  HKG_STAGE_APPROX = 3,   /* exact scan of the best leaf (first bound)       */
  HKG_STAGE_COLLECT = 4,  /* (query, leaf) work list                         */
  HKG_STAGE_REFINE = 5,   /* iSAX or BF16 filter + exact refinement          */
  HKG_STAGE_TOPK = 6,     /* final per-query top-k                           */
  HKG_STAGE_SCAN = 7,     /* PSCAN full scan                                 */
  HKG_STAGE_BOUND = 8,    /* BF16 path: upper-bound histogram pass (k-th
                             distance bound before survivors are collected)  */
  HKG_NUM_STAGES = 9
};

typedef struct hkg_query_stats {
  uint64_t queries;
  uint64_t leaves_total;          /* per query, summed over queries            */
  uint64_t leaves_after_eapca;    /* (query, leaf) pairs whose EAPCA bound is
                                     within the query's k-th distance bound
                                     after the approximate stage, the leaves
                                     that stage scanned included             */
  uint64_t series_total;
  uint64_t series_lb_checked;     /* series whose lower bound was evaluated    */
  uint64_t series_refined;        /* series whose exact distance was computed  */
  uint64_t bytes_algorithmic;     /* HBM bytes the plan had to touch           */
  double device_ms;               /* device time of the query kernels          */
  uint32_t overflow_retries;      /* candidate-buffer overflows re-run         */
  uint32_t launches;              /* kernel launch sets (batches)              */
  uint64_t bf16_pairs;            /* (query, series) pairs the BF16 bound saw  */
  double stage_ms[HKG_NUM_STAGES];         /* with profile=1: summed stage times */
  uint64_t stage_bytes[HKG_NUM_STAGES];    /* algorithmic HBM bytes per stage    */
  uint64_t stage_launches[HKG_NUM_STAGES]; /* kernel launches of each stage      */
  uint64_t mode_leaves_first;       /* modes: leaves scanned before the candidates */
  uint64_t mode_candidates;         /* epsilon search: candidate leaves            */
  uint64_t mode_candidates_scanned; /* epsilon search: candidate leaves examined   */
  uint64_t mode_sequential;         /* epsilon search: queries whose candidates
                                       were scanned in file order (the reference's
                                       sequential branch)                          */
} hkg_query_stats;

typedef struct hkg_index_info {
  uint64_t n_series;
  uint32_t length;
  uint32_t paa_segments;
  uint32_t sax_bit_cardinality;
  uint32_t max_leaf_size;
  uint64_t n_leaves;
  uint64_t n_nodes;
  uint64_t device_bytes;          /* HBM held by this index                    */
  int has_tree;                   /* 0 for hkg_open_raw (scan only)            */
  int has_ids;                    /* an original-id map is attached            */
} hkg_index_info;

/* ---- lifetime ----------------------------------------------------------- */
void hkg_default_options(hkg_options *opt);
void hkg_default_query_opts(hkg_query_opts *q);

/* Load a Hercules index directory (root.idx, leaves_raw.idx,
   leaves_sims.idx, as written with --serial 1 --sims) into HBM as SoA. */
hkg_status hkg_open_dir(const char *index_dir, const hkg_options *opt, hkg_index **out);

/* A flat dataset for the full scan only (no tree); X is host memory,
   row-major n x L float32, L % 8 == 0.
   Series lengths: the query paths take L up to 16384 (the paper's longest);
   the BF16 path takes L = 96, 128 and 256; the device builder (hkg_build*)
   takes L up to 2048 (longer indexes come from the reference's builder and
   hkg_open_dir). */
hkg_status hkg_open_raw(const float *X, uint64_t n, uint32_t L, const hkg_options *opt,
                        hkg_index **out);

/* Attach an id map (leaf-order row -> public id, n distinct 64-bit
   entries; the device ranks them, so ties in d2 still break by public id);
   without one the public id of a series is its leaf-order position
   (file_pos + offset in leaf, the reference's "Pos"). */
hkg_status hkg_set_ids(hkg_index *ix, const uint64_t *ids, uint64_t n);

hkg_status hkg_info(const hkg_index *ix, hkg_index_info *info);
/* Parse and validate an index directory on the host only (no device):
   root.idx structure and the sizes of leaves_raw.idx / leaves_sims.idx. */
hkg_status hkg_inspect_dir(const char *index_dir, hkg_index_info *info);
void hkg_close(hkg_index *ix);

/* ---- queries ------------------------------------------------------------ */
/* Q: nq x L float32 host rows.  k: per-query k (1..256), or NULL to use
   kmax for all.  Outputs ids/d2: nq x kmax, host memory. */
hkg_status hkg_knn(hkg_index *ix, const float *Q, uint32_t nq, const uint32_t *k, uint32_t kmax,
                   const hkg_query_opts *qo, uint64_t *ids, float *d2, hkg_query_stats *stats);

/* Same, with Q, k, ids (uint64) and d2 already in device memory; runs on
   the given HIP stream (cast to void*; NULL = the handle's stream) and
   synchronises it before returning.  Every dk[i] must be in [1, kmax]: the
   device checks them and the call returns HKG_EINVAL otherwise (the
   outputs are then undefined). */
hkg_status hkg_knn_device(hkg_index *ix, const float *dQ, uint32_t nq, const uint32_t *dk,
                          uint32_t kmax, const hkg_query_opts *qo, uint64_t *dids, float *dd2,
                          void *stream, hkg_query_stats *stats);

/* Exact k-NN by the full scan (the kernels of hkg_knn_device with path
   HKG_PATH_SCAN, identical answers) that allocates nothing: no hipMalloc,
   hipFree or allocating library call.  No counterpart in the reference; for
   callers that cannot allocate device memory (code between HIP graph
   captures, tight serving loops, hosts that forbid allocation).  Its scratch
   is a 16 MiB device array inside the library (one per device, held until
   the process ends, set up when an index is opened on the device); a call
   whose scratch would not fit -- beyond 64 queries at kmax 256 on a 256-CU
   MI355X -- returns HKG_EINVAL.  dQ (nq x L), dk (nq, each in [1, kmax]),
   dids and dd2 (nq x kmax) are device memory; runs on `stream` (NULL = the
   handle's stream) and synchronises it before returning.  Calls on one
   device share the array and are serialised by a lock inside the library. */
hkg_status hkg_scan_knn_device_noalloc(hkg_index *ix, const float *dQ, uint32_t nq, const uint32_t *dk,
                                       uint32_t kmax, uint64_t *dids, float *dd2, void *stream);

/* ---- launch-shape autotuner --------------------------------------------- */
/* The reference has no tuner: its query-time shape is the thread counts set
   by hand (--num-threads / --num-read-threads, hercules.c:104-105,227;
   PSCAN's --num-query-threads / --dbsize, parallel_scan.c:64-65).  Here the
   shapes are timed on the index itself, for its series length: the queries
   per device batch of a path and the persistent grid of its main kernel
   (blocks per CU): the iSAX filter + refine kernel on the FP32 index paths,
   the scan kernel on the full scan; on the batched path also its filter
   (per item or leaf-ordered, hkg_query_opts.batched_filter), kept for calls
   that leave batched_filter at 0 and whose batch is the tuned one; on the
   BF16 path the filter's 32-query MFMA tiles per wave
   (hkg_query_opts.bf16_query_tiles), with batches of whole block groups.
   Each shape is timed on the whole batches the sample holds.  Every
   candidate must return the same answers, bit for bit (HKG_EINTERNAL
   otherwise). */
typedef struct hkg_tuning {
  uint32_t path;                 /* the path tuned (hkg_path; AUTO tunes BATCHED) */
  uint32_t length;               /* series length of the index                    */
  uint32_t batch;                /* queries per device batch chosen               */
  uint32_t filter_blocks_per_cu; /* chosen grid of the FP32 filter + refine
                                    (0: the path does not launch it)              */
  uint32_t candidates;           /* shapes timed                                  */
  uint32_t scan_blocks_per_cu;   /* chosen grid of the full scan (scan path; 0
                                    otherwise)                                    */
  uint32_t batched_filter;       /* batched path: the filter chosen (1 per item,
                                    2 leaf-ordered; 0 on the other paths)         */
  uint32_t bf16_query_tiles;     /* BF16 path: query tiles per filter wave chosen
                                    (hkg_query_opts.bf16_query_tiles; 0 on the
                                    other paths)                                  */
  double ms_per_query;           /* device time per query, the chosen shape       */
  double ms_per_query_default;   /* device time per query, the default shape      */
} hkg_tuning;

/* Time the candidate shapes of a path on nq sample queries (host rows, k as
   in hkg_knn), keep the fastest in the handle (of shapes within 1% of the
   fastest, the one with the largest batch) and return it.  Later calls
   on that path with hkg_query_opts.batch = 0 use the tuned batch; the FP32
   index paths and the modes use the tuned grid.  Not to be called while
   other calls on the same handle run. */
hkg_status hkg_autotune(hkg_index *ix, const float *Q, uint32_t nq, const uint32_t *k, uint32_t kmax,
                        uint32_t path, hkg_tuning *out);

/* ---- index build (GPU bulk loader, same on-disk format) ---------------- */
typedef struct hkg_build_opts {
  uint32_t leaf_size;      /* maximum series per leaf                       */
  uint32_t init_segments;  /* initial EAPCA segments at the root (1)        */
  uint32_t paa_segments;   /* iSAX segments (16)                            */
  int device;
} hkg_build_opts;

void hkg_default_build_opts(hkg_build_opts *bo);
/* Build from host rows (n x L float32, original order). */
hkg_status hkg_build(const float *X, uint64_t n, uint32_t L, const hkg_build_opts *bo,
                     hkg_index **out);
/* Build from rows already in device memory (n x L float32, original order);
   the library takes ownership of dX (frees it with hipFree). */
hkg_status hkg_build_device(float *dX, uint64_t n, uint32_t L, const hkg_build_opts *bo,
                            hkg_index **out);
/* Write a resident index in the reference's on-disk format (directory must
   exist).  The id map, if any, is written to ids.u64 beside it. */
hkg_status hkg_write_dir(const hkg_index *ix, const char *index_dir);

/* ---- workload generator (device) ---------------------------------------- */
/* kind 0: z-normalised random walks; kind 1: N(0,1) -> L2 -> z-normalised.
   Writes n x L float32 rows for ids first..first+n-1 into device memory. */
hkg_status hkg_gen_series_device(int kind, uint64_t seed, uint64_t first, uint64_t n, uint32_t L,
                                 float *dX, void *stream);
/* Noisy queries q_i = znorm(X[src_i] + sigma_i N(0,1)), src_i drawn from
   stream i of query_seed; dQ: nq x L device, sigmas/src host. */
hkg_status hkg_gen_noisy_queries_device(int kind, uint64_t data_seed, uint64_t N,
                                        uint64_t query_seed, uint32_t nq, uint32_t L,
                                        const float *sigmas, float *dQ, uint64_t *src_ids);

/* ---- serving loop (continuous batching) --------------------------------- */
typedef struct hkg_server_opts {
  uint32_t max_batch;        /* queries admitted per device batch (1024)     */
  uint64_t scratch_bytes;    /* device scratch budget of all batches in
                                flight; a batch admits queries while its
                                share (scratch_bytes / max_inflight) holds
                                them, each charged the path's scratch at the
                                batch's k class (the smallest power of two
                                covering its requests' k) (0 = 16 GiB)      */
  uint32_t queue_capacity;   /* pending requests before HKG_EBUSY (65536)    */
  uint32_t path;             /* hkg_path of the batches (0 = BATCHED)        */
  uint32_t max_inflight;     /* batches on the device at once (2)            */
  uint32_t max_k;            /* largest k a request may ask for (256)        */
  uint32_t bf16_bound_pass;  /* as in hkg_query_opts                         */
  uint32_t reserved;
} hkg_server_opts;

void hkg_default_server_opts(hkg_server_opts *so);
hkg_status hkg_server_start(hkg_index *ix, const hkg_server_opts *so, hkg_server **out);
/* Submit one query (host row of L floats); returns a ticket. */
hkg_status hkg_submit(hkg_server *s, const float *q, uint32_t k, uint64_t *ticket);
/* Non-blocking: HKG_ENOTREADY until done; then copies k results out and
   reports submit->complete latency in microseconds. */
hkg_status hkg_poll(hkg_server *s, uint64_t ticket, uint64_t *ids, float *d2, double *latency_us);
/* Blocking variant of hkg_poll. */
hkg_status hkg_wait(hkg_server *s, uint64_t ticket, uint64_t *ids, float *d2, double *latency_us);
void hkg_server_stop(hkg_server *s);

/* ---- device memory helpers (for callers without their own HIP binding) -- */
void *hkg_device_alloc(uint64_t bytes);               /* NULL on failure        */
void hkg_device_free(void *p);
/* kind: 1 host->device, 2 device->host, 3 device->device (synchronous) */
hkg_status hkg_memcpy(void *dst, const void *src, uint64_t bytes, int kind);
hkg_status hkg_device_sync(void);

/* ---- misc --------------------------------------------------------------- */
const char *hkg_status_str(hkg_status s);
/* Last error message of the calling thread (static storage). */
const char *hkg_last_error(void);
/* The card-256 iSAX breakpoints used by the device kernels (255 floats). */
void hkg_breakpoints256(float *out255);

#ifdef __cplusplus
}
#endif
#endif /* HKG_H */
