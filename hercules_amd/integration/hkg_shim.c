/* hkg_shim.c -- the reference's query functions on the MI355X engine:
 * exact_de_knn_search_psq_lf_sims_adaptive (exact and epsilon search),
 * exact_ng_knn_search (nprobes), and the exact search of the two siblings the
 * reference dispatches to without --sims (exact_de_knn_search_psq_lf) and
 * with --num-threads 1 (exact_de_knn_search).
 *
 * Compiled with the reference's headers and linked into the reference's own
 * main (hercules.c) in place of the library's definitions of these four
 * functions; every other reference function stays as it is.  The index
 * directory is the one the reference loaded (index->settings->root_directory,
 * written with --serial 1 --sims); the engine uploads it to HBM once, on the
 * first query. */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "hercules_query_engine.h"   /* reference: query_result, query_settings, stats calls */
#include "hkg.h"

static hkg_index *g_hkg = NULL;

static hkg_index *hkg_for(struct hercules_index *index)
{
  if (!g_hkg) {
    hkg_options o;
    hkg_default_options(&o);
    if (hkg_open_dir(index->settings->root_directory, &o, &g_hkg) != HKG_OK) {
      fprintf(stderr, "hkg: %s\n", hkg_last_error());
      exit(-1);                      /* the reference's convention for fatal errors */
    }
  }
  return g_hkg;
}

/* one query through the engine, then the reference's records rank by rank
 * (hercules_query_engine.c:1062-1073, :2786-2797); the records print
 * sqrtf(distance), so the squared distance goes in as the reference keeps it */
static void hkg_answer(struct hercules_index *index, ts_type *query_ts, unsigned int k, unsigned int q_id,
                       char *qfilename, const hkg_query_opts *qo)
{
  uint64_t *ids = malloc(sizeof(uint64_t) * k);
  float *d2 = malloc(sizeof(float) * k);
  if (hkg_knn(hkg_for(index), query_ts, 1, NULL, k, qo, ids, d2, NULL) != HKG_OK) {
    fprintf(stderr, "hkg: %s\n", hkg_last_error());
    exit(-1);
  }
  for (unsigned int pos = 0; pos < k; ++pos) {
    struct query_result r;
    r.distance = d2[pos];            /* FLT_MAX for ranks past the number of distinct distances */
    r.node = NULL;
    r.max_distance = FLT_MAX;
    r.pqueue_position = 0;
    update_query_stats(index, q_id, pos + 1, r);
    get_query_stats(index, pos + 1);
    print_query_stats(index, q_id, pos + 1, qfilename);
  }
  fflush(stdout);
  free(ids);
  free(d2);
}

/* hercules_query_engine.c:706-1083 (include/hercules_query_engine.h:284-289):
   exact search, or the epsilon search when epsilon > 0 */
void exact_de_knn_search_psq_lf_sims_adaptive(ts_type *query_ts, ts_type *query_ts_reordered, int *query_order,
                                              unsigned int offset, ts_type *query_paa,
                                              struct hercules_index *index, ts_type minimum_distance,
                                              ts_type epsilon, double delta, unsigned int k, unsigned int q_id,
                                              char *qfilename, query_settings q_settings)
{
  (void)query_ts_reordered; (void)query_order; (void)offset; (void)query_paa; (void)minimum_distance;
  hkg_query_opts qo;
  hkg_default_query_opts(&qo);
  qo.path = HKG_PATH_INDEX;          /* one query: the per-query index path */
  qo.epsilon = epsilon;              /* 0: exact; > 0: the reference's epsilon search */
  qo.delta = (float)delta;           /* accepted; the reference ignores it too */
  qo.approx_leaves = (uint32_t)q_settings.approx_stop_condition;
  qo.eapca_threshold = q_settings.eapca_threshold;
  qo.sax_threshold = q_settings.sax_threshold;
  hkg_answer(index, query_ts, k, q_id, qfilename, &qo);
}

/* hercules_query_engine.c:2623-2799 (include/hercules_query_engine.h:209-212):
   the no-guarantee search over nprobes leaves */
void exact_ng_knn_search(ts_type *query_ts, ts_type *query_ts_reordered, int *query_order, unsigned int offset,
                         struct hercules_index *index, ts_type minimum_distance, unsigned int k, unsigned int q_id,
                         char *qfilename, unsigned int nprobes, query_settings q_settings)
{
  (void)query_ts_reordered; (void)query_order; (void)offset; (void)minimum_distance; (void)q_settings;
  hkg_query_opts qo;
  hkg_default_query_opts(&qo);
  qo.nprobes = nprobes;
  hkg_answer(index, query_ts, k, q_id, qfilename, &qo);
}

/* The siblings answer the same exact k-NN; their epsilon searches traverse
   the tree in their own orders (approximate_knn_search, :176-222), which the
   engine does not replay, so epsilon > 0 is refused rather than answered on
   the CPU or with another traversal's answers. */
static void hkg_exact_only(const char *fn, ts_type epsilon)
{
  if (epsilon != 0.0f) {
    fprintf(stderr, "hkg: %s: --epsilon %g is not supported on this path; run with --sims and more than one "
                    "thread (the epsilon search of exact_de_knn_search_psq_lf_sims_adaptive)\n",
            fn, (double)epsilon);
    exit(-1);                        /* the reference's convention for fatal errors */
  }
}

/* hercules_query_engine.c:2215-2448 (include/hercules_query_engine.h:203-207):
   the serial exact search, called for --num-threads 1
   (hercules_file_loaders.c:239-244) */
void exact_de_knn_search(ts_type *query_ts, ts_type *query_ts_reordered, int *query_order, unsigned int offset,
                         struct hercules_index *index, ts_type minimum_distance, ts_type epsilon, ts_type r_delta,
                         unsigned int k, unsigned int q_id, char *qfilename, query_settings q_settings)
{
  (void)query_ts_reordered; (void)query_order; (void)offset; (void)minimum_distance; (void)r_delta;
  (void)q_settings;
  hkg_exact_only("exact_de_knn_search", epsilon);
  hkg_query_opts qo;
  hkg_default_query_opts(&qo);
  qo.path = HKG_PATH_INDEX;
  hkg_answer(index, query_ts, k, q_id, qfilename, &qo);
}

/* hercules_query_engine.c:1398-1690 (include/hercules_query_engine.h:256-261):
   the multithreaded exact search without the SIMS cache, called without
   --sims (hercules_file_loaders.c:258-262) */
void exact_de_knn_search_psq_lf(ts_type *query_ts, ts_type *query_ts_reordered, int *query_order,
                                unsigned int offset, struct hercules_index *index, ts_type minimum_distance,
                                ts_type epsilon, double delta, unsigned int k, unsigned int q_id, char *qfilename,
                                query_settings q_settings)
{
  (void)query_ts_reordered; (void)query_order; (void)offset; (void)minimum_distance; (void)delta;
  (void)q_settings;
  hkg_exact_only("exact_de_knn_search_psq_lf", epsilon);
  hkg_query_opts qo;
  hkg_default_query_opts(&qo);
  qo.path = HKG_PATH_INDEX;
  hkg_answer(index, query_ts, k, q_id, qfilename, &qo);
}
