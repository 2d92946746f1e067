/*
 * hera_em_gen.h -- seeded synthetic equivalence-class tables.
 *
 * No counterpart in the reference (hera builds classes from alignments,
 * src/hash_align.c:1564-1602).  The generator reproduces hera's layout rules
 * so benchmarks and parity tests see hera-shaped input:
 *   - transcripts in gene clusters of uniform size [cluster_min, cluster_max]
 *     with contiguous ids;
 *   - integer lengths, lognormal(len_mu, len_sigma) clamped and rounded
 *     (REF_INF->len is u32, src/hash_align.h:116);
 *   - class sizes geometric(geom_p) on {1, 2, ...} capped at cls_max and at
 *     the cluster size; members distinct, drawn within one uniformly chosen
 *     cluster, ascending;
 *   - class counts Zipf(zipf_s) over a seeded permutation of ranks, scaled to
 *     n_frag, rounded to nearest with a minimum of 1;
 *   - single-member classes folded into the unique counts u_t, as hera does
 *     (src/hash_align.c:1571-1572), so the class table holds only n >= 2.
 */
#ifndef HERA_EM_GEN_H_
#define HERA_EM_GEN_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
	uint32_t n_tx;        /* transcripts                                  */
	uint32_t n_cls_all;   /* classes drawn, singletons included            */
	uint64_t n_frag;      /* fragment total the Zipf counts are scaled to  */
	uint32_t cluster_min, cluster_max;
	double   len_mu, len_sigma;
	uint32_t len_min, len_max;
	double   geom_p;
	uint32_t cls_max;
	double   zipf_s;
	uint64_t seed;
} hem_gen_params;

typedef struct {
	uint32_t  n_tx, n_cls;   /* n_cls = multi-member classes               */
	uint64_t  nnz;
	uint32_t *cls_off;       /* n_cls + 1                                  */
	uint32_t *cls_ref;       /* nnz                                        */
	double   *cls_count;     /* n_cls                                      */
	uint32_t *tx_len;        /* n_tx                                       */
	uint32_t *unique;        /* n_tx: singleton-class fragments            */
	uint32_t  n_clusters;
	uint32_t *cluster_off;   /* n_clusters + 1                             */
	uint64_t  frag_total;    /* sum of all class counts produced           */
} hem_gen_problem;

/* Defaults of the benchmark generator for the given sizes: clusters 2-16,
 * lognormal(7.3, 0.8) lengths in [200, 20000], geometric(0.35) sizes capped
 * at 32, Zipf(1.1) counts. */
void hem_gen_default(hem_gen_params *g, uint32_t n_tx, uint32_t n_cls_all,
                     uint64_t n_frag, uint64_t seed);
int hem_generate(const hem_gen_params *g, hem_gen_problem **out);
void hem_gen_free(hem_gen_problem *p);

#ifdef __cplusplus
}
#endif
#endif
