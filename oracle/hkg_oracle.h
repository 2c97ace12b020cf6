/*
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing measured or shipped.
 *
 * A plain-C restatement of the reference's arithmetic on the exact k-NN
 * path, each function citing the reference lines it follows, plus the
 * workload generator specification shared with the product's HIP generator.
 */
#ifndef HKG_ORACLE_H
#define HKG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- workload generator (specification; the HIP generator must match) -- */
uint64_t orc_mix64(uint64_t x);
uint64_t orc_draw(uint64_t seed, uint64_t stream, uint64_t j);
double   orc_uniform(uint64_t bits);
double   orc_log(double u);
double   orc_crsqrt(double x);
void     orc_normal_pair(uint64_t seed, uint64_t stream, uint64_t pair, double *z0, double *z1);
/* kind 0: z-normalised random walk; kind 1: N(0,1) -> L2 -> z-normalised */
void orc_gen_series(int kind, uint64_t seed, uint64_t first_id, uint64_t n, uint32_t L,
                    float *out, int nthreads);
/* noisy queries: q_i = znorm(X[src_i] + sigma_i * N(0,1)); src_i drawn
   uniformly in [0, N) from stream i of query_seed (or src_ids given) */
void orc_gen_noisy_queries(int kind, uint64_t data_seed, uint64_t N, uint64_t query_seed,
                           uint32_t nq, uint32_t L, const float *sigmas,
                           float *out, uint64_t *src_ids);

/* ---- reference arithmetic (bit-exact restatements) ---------------------- */
float orc_ed2(const float *q, const float *x, int L, float bound);            /* ts.c:83-109 */
void  orc_paa(const float *x, int L, int w, float *out);                       /* sax.c:46-56 */
void  orc_breakpoints256(float *out255);                                       /* closed form */
void  orc_sax_word(const float *paa, int w, const float *bp, uint8_t *out);    /* sax.c:86-135 */
float orc_lb_sax(const float *paa, const uint8_t *sax, int L, const float *bp);/* sax.c:310-479 */
void  orc_seg_mean_std(const float *x, int s, int e, float *mu, float *sd);    /* calc_utils.c:48-112 */
float orc_lb_eapca(const float *q, int nseg, const int16_t *seg_end,
                   const float *ind, int use_fma);                             /* hercules_node.c:879-924 */

/* ---- answers ------------------------------------------------------------ */
/* k smallest distinct FP32 d2 per query, smallest id among equal d2;
   unfilled ranks get d2 = FLT_MAX and id = UINT64_MAX. */
int orc_knn_bruteforce(const float *X, uint64_t N, int L, const float *Q, uint32_t nq,
                       uint32_t k, uint64_t *ids, float *d2, int nthreads);
/* FP64 squared distance (error reporting) */
double orc_ed2_f64(const float *q, const float *x, int L);

/* ---- approximate query modes on an index tree -------------------------- */
/* Node table in pre-order, ORC_NODE_FIELDS int64 per node:
     {is_leaf, left, right, seg_off, nseg, split_from, split_to, split_idx,
      file_pos, size}
   seg_end/ind: every node's segments at seg_off (ind: 4 per segment, {max
   mean, min mean, max sd, min sd}); split_value per node; X and sax (16
   bytes per row) in leaf-file order, so a row's id is its position in
   leaves_raw.idx.
   nprobes > 0: exact_ng_knn_search (hercules_query_engine.c:2623-2799);
   else the epsilon search of exact_de_knn_search_psq_lf_sims_adaptive
   (:706-1083) with approximate_greedy_knn_search's leaf budget
   approx_budget (--approx-stop-condition) and the EAPCA / SAX thresholds
   (--eapca-threshold, --sax-threshold).  info[4 per query]: leaves scanned
   before the candidate stage, candidate leaves, the candidates' branch (0
   none, 1 sequential, 2 iSAX + refinement), candidate leaves scanned. */
#define ORC_NODE_FIELDS 10
int orc_knn_mode(const int64_t *nodes, const float *split_value, uint32_t nnodes, const int16_t *seg_end,
                 const float *ind, const float *X, const uint8_t *sax, int L, const float *Q, uint32_t nq,
                 uint32_t k, float epsilon, int approx_budget, uint32_t nprobes, float eapca_th, float sax_th,
                 uint64_t *ids, float *d2, uint32_t *info, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
