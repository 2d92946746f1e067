/*
 * em_oracle.h -- CPU restatement of hera's EM (TEST INFRASTRUCTURE ONLY).
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker; never linked by the product library.
 */
#ifndef EM_ORACLE_H_
#define EM_ORACLE_H_
#include <stdint.h>

typedef struct {
	uint32_t min_round;   /* src/EM.c:325 */
	uint32_t max_round;   /* src/EM.c:326 */
	double stop_min;      /* src/EM.c:292 */
	double stop_rel;      /* src/EM.c:295 */
	int step_form;        /* 0 sample sd (src/EM.c:338), 1 L2 (hera-EM) */
	int clamp_form;       /* 0 zero (src/EM.c:351), 1 u2/10 (hera-EM)   */
	int alpha_clamp;      /* 1: alpha <= -1 (src/EM.c:346)               */
	int fallback;         /* 1: monotonicity fallback (no reference)     */
} oracle_opts;

typedef struct {
	uint32_t cap;         /* capacity of alpha / stop                    */
	uint32_t k;           /* iterations run                              */
	double *alpha;        /* step length per iteration                   */
	uint32_t *stop;       /* stop flag after each iteration's 3rd map    */
	double *last_u;       /* n_tx: extrapolated point of the last
	                         iteration (the input of its 3rd map), or NULL */
	uint32_t *fallback;   /* per iteration: 1 if the fallback kept u2, or
	                         NULL                                        */
} oracle_trace;

void oracle_opts_default(oracle_opts *o);
double oracle_norm_sd(const double *arr, uint32_t n);
int oracle_estimate_count(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                          const uint32_t *ref, const double *cnt,
                          const uint32_t *len, const uint32_t *count,
                          const oracle_opts *opt, const double *init,
                          double *tpm, double *est, double *lambda,
                          oracle_trace *tr);
int oracle_em_map(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                  const uint32_t *ref, const double *cnt, const uint32_t *len,
                  const uint32_t *count, const double *lam_in,
                  double *lam_out, uint32_t *stop);
/* The same two with per-member weights w (CSR order; NULL = unweighted):
 * D_c = sum_k lam_{t_k} w_k and each member adds lam_t w_k n_c / D_c. */
int oracle_estimate_count_w(uint32_t n_tx, uint32_t n_cls,
                            const uint32_t *off, const uint32_t *ref,
                            const double *cnt, const double *w,
                            const uint32_t *len, const uint32_t *count,
                            const oracle_opts *opt, const double *init,
                            double *tpm, double *est, double *lambda,
                            oracle_trace *tr);
int oracle_em_map_w(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                    const uint32_t *ref, const double *cnt, const double *w,
                    const uint32_t *len, const uint32_t *count,
                    const double *lam_in, double *lam_out, uint32_t *stop);
double oracle_loglik(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                     const uint32_t *ref, const double *cnt,
                     const uint32_t *len, const uint32_t *count,
                     const double *lam);
double oracle_grad(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                   const uint32_t *ref, const double *cnt,
                   const uint32_t *count, const double *theta,
                   const double *ell, double *g_theta, double *g_ell);
void oracle_hvp(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                const uint32_t *ref, const double *cnt,
                const uint32_t *count, const double *theta, const double *ell,
                const double *vec, double *hv);
int oracle_bootstrap_batch(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                           const uint32_t *ref, const uint32_t *len,
                           uint32_t n_rep, const uint32_t *counts,
                           const double *cls_counts, const oracle_opts *opt,
                           uint32_t n_thread, double *est, uint32_t *k_out);
double oracle_now(void);
/* ell(lam) - ell(mu) of oracle_loglik's objective (w: member weights or
 * NULL), summed as per-term differences (log1p of relative changes);
 * *scale (if not NULL) = the sum of the terms' magnitudes.  The fallback
 * keeps u2 when ell(lam) - ell(u2) < -1e-10 * scale. */
double oracle_dloglik(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                      const uint32_t *ref, const double *cnt, const double *w,
                      const uint32_t *len, const uint32_t *count,
                      const double *lam, const double *mu, double *scale);


/* bootstrap resampling restatement */
void oracle_philox(uint64_t key, const uint32_t ctr[4], uint32_t out[4]);
uint32_t oracle_resample_plan(const uint32_t *w, uint32_t K, uint32_t N,
                              uint32_t *start, uint32_t *cw);
int oracle_resample(const uint32_t *w, uint32_t K, uint64_t seed,
                    uint32_t *out);
int oracle_bootstrap_resampled(uint32_t n_tx, uint32_t n_cls,
                               const uint32_t *off, const uint32_t *ref,
                               const uint32_t *len, const uint32_t *w,
                               uint32_t n_rep, const uint64_t *seeds,
                               const oracle_opts *opt, uint32_t n_thread,
                               double *est, uint32_t *k_out);
int oracle_write_abundance(const char *path, uint32_t n, const char *names,
                           uint32_t name_len, const uint32_t *count,
                           const uint32_t *len, const double *eff_len,
                           const double *overlap, const double *grama);
int oracle_write_gene_abundance(const char *path, uint32_t n,
                                const char *names, uint32_t name_len,
                                const uint32_t *gene, const char *gene_name,
                                uint32_t l_gene, const double *overlap,
                                const double *grama);

/* hera-EM's simplified-model SQUAREM (restated in em_oracle.c) */
typedef struct {
	uint32_t noise_tx;    /* the noise component's transcript, or >= n_tx */
	uint32_t max_iter;
	int locking;          /* lock entries after each extrapolation        */
	int filtering;        /* filter on the second EM of each loop         */
	double lock_thr;      /* 1e-5  LOCK_THRESHOLD (hera-EM/utility.h:13)  */
	double filter_thr;    /* 1e-6  FILTERING_THRESHOLD (utility.h:14)     */
	double stop_rel;      /* 1e-3  THRESHOLD1 (utility.h:10)              */
	double stop_min;      /* 1e-7  THRESHOLD2 (utility.h:11)              */
} oracle_he_opts;
typedef struct {
	uint32_t iters, converged, n_locked, removed_members, removed_classes;
} oracle_he_stats;
int oracle_hera_em(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                   const uint32_t *ref, const double *cnt, const double *w,
                   const oracle_he_opts *o, const double *p0, double *p_out,
                   oracle_he_stats *st);

uint32_t oracle_rand_r(uint32_t *seed);
int oracle_hera_sample(const uint32_t *unique, uint32_t n_tx, uint32_t *seed,
                       uint32_t *count);

#endif
