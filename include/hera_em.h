/*
 * hera_em.h -- C ABI of the MI355X SQUAREM EM abundance estimator.
 *
 * Plain C: pointers, sizes and int status codes; no HIP or torch types.
 * All entry points are thread-safe with respect to distinct hem_problem
 * handles; calls on one handle are serialised internally.
 *
 * Reference interfaces replaced (citations into bioturing/hera):
 *   estimate_count  src/EM.h:28, defined src/EM.c:315-387 (declared in
 *                   hera_dropin.h, which reads hera's globals)
 *   M_step/E_step   src/EM.c:271-313  -> device kernels driven by hem_solve
 *   Norm            src/hash_align.c:150-164 -> FP64 device reduction
 *   bootstrap loop  src/EM.c:123-150, 233-261 -> hem_bootstrap
 */
#ifndef HERA_EM_H_
#define HERA_EM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define HEM_OK             0
#define HEM_ERR_INVALID    1  /* bad argument or malformed class table      */
#define HEM_ERR_NOMEM      2  /* host or device allocation failed           */
#define HEM_ERR_DEVICE     3  /* HIP runtime error                          */
#define HEM_ERR_NODEVICE   4  /* no usable gfx950 device                    */
#define HEM_ERR_RANGE      5  /* size beyond a documented limit             */
#define HEM_ERR_IO         6  /* a file could not be opened or written      */

/* Last error message of the calling thread (static storage, never NULL). */
const char *hem_last_error(void);

/* ---- problem description ----------------------------------------------- */

/* The class table in CSR form.  Equivalent to hera's Gclass + Ref_inf:
 * class c has members cls_ref[cls_off[c] .. cls_off[c+1]) (u32 transcript
 * ids < n_tx, duplicates allowed, any order) and cls_count[c] fragments.
 * tx_len[t] is REF_INF->len[t] (must be > 0).
 *
 * member_weight (NULL in hera's model): a weight w_k >= 0 per member, as
 * hera-EM's alignment probabilities (c_prob, hera-EM/smodel.h:15-20).  The
 * map becomes D_c = sum_k lam_{t_k} w_k (member order) and, for each
 * member, ((lam_t w_k) n_c) / D_c into overlap_t; w = 1 is hera's map bit
 * for bit.  With unit lengths and the noise as one more transcript that is
 * a member of every class with weight 4^-read_length, this is hera-EM's
 * E-step and renormalisation (hera-EM/em.c:75-157, transdis.c:54-61) up
 * to the scale of lam; locking and filtering are not part of it.
 * Weighted tables run plain FP64 solves (hem_solve) only. */
typedef struct {
	uint32_t        n_tx;
	uint32_t        n_cls;
	const uint32_t *cls_off;    /* n_cls + 1 */
	const uint32_t *cls_ref;    /* cls_off[n_cls] */
	const double   *cls_count;  /* n_cls */
	const uint32_t *tx_len;     /* n_tx */
	const double   *member_weight; /* cls_off[n_cls], or NULL */
} hem_classes;

/* Bulk E-step precision.  FP64 (default): the class and transcript passes
 * add in the reference's order (members in class order, classes in
 * ascending id per transcript, (lam_t * n_c) / D_c per term).  FP32 (the
 * bulk form): on the component tiles, class sums of the densities rounded
 * to FP32, one FP32 n_c / D_c per class and FP32 terms, each transcript's
 * terms summed in FP64 (tables without member weights); on the two global
 * passes (a component over the tile limits), gathered densities, class
 * denominators and per-transcript sums in FP32.  The M-step, SQUAREM
 * state, norms, stop test and log-likelihood stay FP64 in both. */
#define HEM_PREC_FP32 0
#define HEM_PREC_FP64 1

/* SQUAREM step length.  SD: alpha = -sd(r)/sd(v) with the sample standard
 * deviation, as src/EM.c:338-339 with Norm (src/hash_align.c:150-164).
 * L2: alpha = -||r||/||v|| (hera-EM/utility.c:16-41). */
#define HEM_STEP_SD 0
#define HEM_STEP_L2 1

/* SQUAREM norm reductions.  EXACT: the left-to-right sums of Norm
 * (src/hash_align.c:150-164) and of the finalisation (src/EM.c:362-363),
 * bitwise, so with HEM_PREC_FP64 every iterate is bitwise the reference's:
 * each chunk of values becomes a short program (runs whose partial sums
 * stay in one binade, as exact integer sums on its grid, and the other
 * values one by one) and one wave walks the programs in order.  EXACT_WIN:
 * the same bits by an earlier method (chunks run from windows of
 * speculated start values, an in-order lookup, and a chunk whose start
 * falls outside its window added again from the exact start).  SEQ: the
 * same sums, each run by one device lane (slow; kept as the plain
 * restatement).  TREE: fixed-shape parallel sums (bitwise reproducible run
 * to run, ~1 ulp from the reference's). */
#define HEM_NORM_EXACT     0
#define HEM_NORM_TREE      1
#define HEM_NORM_SEQ       2
#define HEM_NORM_EXACT_WIN 3

/* Treatment of extrapolated values.  ZERO: negatives become 0
 * (src/EM.c:351).  DIV10: values <= 1e-300 become u2/10
 * (hera-EM/utility.c:50-53). */
#define HEM_CLAMP_ZERO  0
#define HEM_CLAMP_DIV10 1

typedef struct {
	int      precision;       /* HEM_PREC_*; default FP64                  */
	int      step;            /* HEM_STEP_*; default SD                    */
	int      clamp;           /* HEM_CLAMP_*; default ZERO                 */
	int      alpha_clamp;     /* 1: alpha = min(alpha, -1) (src/EM.c:346)  */
	int      norm;            /* HEM_NORM_*; default EXACT                 */
	uint32_t min_iter;        /* 50   (src/EM.c:325)                       */
	uint32_t max_iter;        /* 1000 (src/EM.c:326)                       */
	double   stop_min_value;  /* 1e-3 (src/EM.c:292)                       */
	double   stop_rel_change; /* 0.01 (src/EM.c:295); e.g. 1e-8 for a
	                             tolerance-mode run                        */
	int      fallback;        /* 0 (default, the reference's loop) or 1:
	                             monotonicity fallback, see below         */
	int      verbose;         /* 1: print hera's "Finish EM" line          */
} hem_opts;

/* Monotonicity fallback (fallback = 1; neither reference has one, nor a
 * likelihood).  Each SQUAREM iteration computes u1, u2, the extrapolated
 * point and lam = M(u') as src/EM.c:329-355 does, then compares the
 * log-likelihood of hem_loglik at lam and at u2: when ell(lam) - ell(u2) <
 * -1e-10 S, with the difference summed as per-term differences (n_c
 * log1p((D_c(lam) - D_c(u2)) / D_c(u2)) per class, u_t log1p((lam_t -
 * u2_t) / u2_t) and -len_t (lam_t - u2_t) per transcript) and S the sum of
 * those terms' magnitudes, the iteration ends at u2 (two plain maps, which
 * never lower the likelihood) and its stop test is the second map's (u2
 * against u1).  The likelihood thus never decreases from one iteration to
 * the next beyond rounding.  hem_stats.fallbacks counts such iterations.
 * With fallback = 0 nothing changes.  hem_solve runs it on the component
 * tiles only (FP64, exact or tree norms, every component within the tile
 * limits; HEM_ERR_RANGE otherwise); member weights enter D_c. */

/* Fills the reference defaults (src/EM.c:324-346, 292-295). */
void hem_opts_default(hem_opts *o);

typedef struct {
	uint32_t iters;        /* SQUAREM iterations k (hera's "rounds")      */
	uint32_t maps;         /* EM map applications (3 per iteration)      */
	uint32_t converged;    /* 1 if the stop rule ended the loop          */
	uint32_t fallbacks;    /* iterations where the fallback fired        */
	double   loglik;       /* FP64 log-likelihood of the final densities
	                          (NaN for a table with member weights)   */
	double   last_alpha;   /* step length of the last iteration          */
	double   ms_solve;     /* device time of the iteration loop          */
	double   ms_total;     /* host wall time of the call                 */
	uint32_t sum_reruns;   /* HEM_NORM_EXACT: chunks of the exact sums
	                          added again from their exact start          */
} hem_stats;

/* ---- device-resident problem ------------------------------------------- */

typedef struct hem_problem hem_problem;

/* Validates the class table, builds the transposed (per-transcript) index
 * and uploads both to the current device.  The caller's arrays are not
 * retained.  Returns HEM_ERR_INVALID for ids >= n_tx, zero lengths,
 * non-monotone offsets or negative / non-finite counts. */
int hem_problem_create(const hem_classes *cls, hem_problem **out);
int hem_problem_destroy(hem_problem *p);

/* Ragged multi-sample handle (no counterpart in the reference, which
 * quantifies one sample per run): n_samples independent class tables side
 * by side in one device problem -- sample q's transcripts are ids
 * tx0_q .. tx0_q + n_tx_q - 1 of the handle (tx0_q = the sum of the earlier
 * samples' n_tx), its classes likewise -- tiled so that no tile holds two
 * samples.  On such a handle hem_solve_batch, hem_bootstrap and
 * hem_bootstrap_counts run n_samples x n_rep EM instances together, each
 * its own EM over its sample's transcripts (uniform start 1/n_tx_q, norms,
 * step, stop test and iteration count per instance), exactly as each
 * sample's own handle would; seeds and iteration counts are per instance,
 * index q * n_rep + r; count and output rows span all samples' transcripts
 * (tpm sums to 1e6 within each sample).  Every component must fit a tile;
 * HEM_NORM_EXACT runs the single-lane sums on such handles (same results).
 * The single-problem entry points (hem_solve, hem_loglik,
 * hem_profile_solve, hem_grad, hem_hvp, hem_stderr) reject it. */
int hem_samples_create(const hem_classes *samples, uint32_t n_samples,
                       hem_problem **out);

/* Replaces the class counts (n_cls values, same validation as creation);
 * the structure stays on the device. */
int hem_problem_set_counts(hem_problem *p, const double *cls_count);

/* Sizes of the device layout (for the caller's bookkeeping and bytes
 * accounting): n_tx, n_cls, nnz, device bytes held. */
int hem_problem_info(const hem_problem *p, uint32_t *n_tx, uint32_t *n_cls,
                     uint64_t *nnz, uint64_t *device_bytes);

/* One SQUAREM EM solve (src/EM.c:315-387).  unique[t] is the unique-read
 * count of transcript t (REF_INF->count).  Outputs (any may be NULL):
 * tpm[t] (hera's grama after finalisation, sums to 1e6), est[t]
 * (est_counts = lambda_t * len_t), lambda[t] (the final per-base density
 * before finalisation). */
int hem_solve(hem_problem *p, const uint32_t *unique, const hem_opts *o,
              double *tpm, double *est, double *lambda, hem_stats *st);

/* FP64 log-likelihood of densities lambda (n_tx values, host memory):
 *   l = sum_t u_t log lambda_t + sum_{c: D_c>0} n_c log D_c
 *       - sum_t len_t lambda_t,
 *   D_c = sum_{t in c} lambda_t, with 0 log 0 = 0.  Classes whose members
 *   are all zero are skipped, as the map skips them (src/EM.c:279-280).
 * hera's map (src/EM.c:271-302) is the EM map of this objective. */
int hem_loglik(hem_problem *p, const uint32_t *unique, const double *lambda,
               double *loglik);

/* ---- batches: many EM instances over one class table --------------------
 * Replicates share the device class table and run together, R lanes per
 * component tile ([transcript][replicate] layouts, replicate fastest);
 * each keeps its own iteration count, step length and stop test, exactly as
 * R separate hem_solve calls would: bitwise so with HEM_NORM_EXACT in
 * either precision (the FP32 bulk form is the arithmetic of hem_solve's
 * tile kernels) and with HEM_NORM_SEQ in FP64.
 * They need every connected component of the class graph to fit a tile
 * (hem_problem_tiles reports n_tiles = 0 otherwise); class counts must be
 * integers below 2^32. */
typedef struct {
	uint32_t n_rep;
	uint32_t iters_min, iters_max;  /* over replicates                   */
	uint32_t n_converged;           /* replicates whose stop rule fired  */
	uint32_t chunk;                 /* replicates per workgroup          */
	uint32_t n_tiles;
	double   ms_resample;           /* device time of the resampling     */
	double   ms_solve;              /* device time of the iteration loop */
	double   ms_total;              /* host wall time of the call        */
	uint32_t sum_reruns;            /* HEM_NORM_EXACT: chunks of the exact
	                                   sums added again (all replicates) */
} hem_batch_stats;

/* Tiling of the class graph used by the batched and standard-error kernels
 * (connected components packed into tiles of about 24 transcripts):
 * number of tiles, largest tile in transcripts and classes.  n_tiles is 0
 * when a component exceeds the tile limits (64 transcripts or 320
 * classes) and for tables with member weights; batches then refuse the
 * table and hem_solve uses its two global passes.  (hem_solve's tile
 * kernels use a second packing of the same components into larger
 * tiles.)  Any output pointer may be NULL. */
int hem_problem_tiles(const hem_problem *p, uint32_t *n_tiles,
                      uint32_t *max_tx, uint32_t *max_cls);
/* The same tiling for a class table on the host, without a device (all
 * three outputs required).  Returns HEM_OK with *n_tiles = 0 when a
 * component exceeds the tile limits. */
int hem_classes_tiling(const hem_classes *cls, uint32_t *n_tiles,
                       uint32_t *max_tx, uint32_t *max_cls);

/* ---- tile configuration and the autotuner --------------------------------
 * No counterpart in the reference (a CPU loop has no tiles).  The packing
 * targets of the two tilings and the batched kernels' replicate chunk.
 * Results do not depend on them in the modes whose sums follow the
 * reference's order (HEM_NORM_EXACT / HEM_NORM_SEQ, either precision):
 * every class and transcript sum runs in the order its list gives,
 * whichever tile holds it.  The tree norms' grouping follows the tiles, and
 * so do their last bits. */
typedef struct {
	uint32_t ptile_tx, ptile_cls, ptile_nnz; /* single-problem tile kernels:
	                                            64 / 320 / 1024           */
	uint32_t tile_tx, tile_cls, tile_nnz;    /* batched and standard-error
	                                            kernels: 24 / 120 / 576   */
	uint32_t chunk;  /* replicates per batched workgroup (1, 2, 4, ...,
	                    32): the most rows of one entity a wave serves at
	                    once is 64 / chunk; 0 = the largest chunk whose
	                    rows fit the LDS budget (default)              */
} hem_tile_config;

/* The defaults above. */
void hem_tile_config_default(hem_tile_config *cfg);
/* hem_problem_create with a tile configuration (NULL: the defaults).
 * Targets are packing targets: a component larger than a target but within
 * the tile limits (64 transcripts, 320 classes) gets a tile of its own.
 * HEM_ERR_INVALID for a target of 0, ptile_tx or tile_tx above 64,
 * ptile_cls or tile_cls above 320, or a chunk that is not 0 or a power of
 * two up to 32.  A chunk larger than fits the LDS budget is lowered to
 * the largest that fits. */
int hem_problem_create_config(const hem_classes *cls,
                              const hem_tile_config *cfg, hem_problem **out);
/* The configuration a handle was created with. */
int hem_problem_get_config(const hem_problem *p, hem_tile_config *cfg);

typedef struct {
	uint32_t n_rep;       /* replicates of the batched workload to tune
	                         for; 0: tune the single-problem tiles only   */
	uint32_t iters;       /* SQUAREM iterations per timed run (default 20
	                         when 0); every run does exactly this many    */
	uint32_t repeats;     /* timed runs per candidate, fastest kept
	                         (default 3 when 0)                           */
	const char *cache_path; /* text file of tuned configurations, or NULL;
	                         read first, appended to after a search       */
} hem_tune_opts;

typedef struct {
	int      from_cache;       /* 1: the configuration came from the file */
	uint32_t n_candidates;     /* configurations timed                    */
	double   us_plain_default; /* device time per SQUAREM iteration of    */
	double   us_plain_best;    /* a plain exact solve (default, chosen)   */
	double   us_batch_default; /* device time per replicate-iteration of  */
	double   us_batch_best;    /* the batched exact solve                 */
	char     key[160];         /* the cache key (T, C, Z, R, device arch) */
} hem_tune_report;

/* Picks the tile configuration for a class table by timing candidates on
 * the current device: the single-problem tile targets (transcripts and
 * members per tile) on plain exact solves of unique, and, with n_rep > 0,
 * the batched tile targets and replicate chunk on an exact batch of n_rep
 * replicates of the table's own counts.  Each candidate is a temporary
 * handle; the defaults are among the candidates and stay the choice
 * unless a candidate is at least 2% faster (the timed runs are short).  With cache_path, a configuration stored
 * under the same key (n_tx, n_cls, members, n_rep, device arch) is
 * returned without timing, and a new search appends one line to the file
 * (concurrent tuners of one key may both append; the last line read
 * wins).  rep may be NULL. */
int hem_autotune(const hem_classes *cls, const uint32_t *unique,
                 const hem_tune_opts *o, hem_tile_config *cfg,
                 hem_tune_report *rep);
/* Host only: the cache key of a table and replicate count on a device
 * arch, and a lookup of it in a cache file (HEM_ERR_INVALID when absent or
 * malformed; the last matching line wins). */
int hem_tune_key(uint32_t n_tx, uint32_t n_cls, uint64_t nnz,
                 uint32_t n_rep, const char *arch, char *key, size_t len);
int hem_tune_cache_lookup(const char *path, const char *key,
                          hem_tile_config *cfg);
int hem_tune_cache_store(const char *path, const char *key,
                         const hem_tile_config *cfg);

/* n_rep EM solves with per-replicate counts: unique[r*n_tx + t] and
 * cls_count[r*n_cls + c] (NULL: the table's own class counts for every
 * replicate).  Outputs (any may be NULL): est[r*n_tx + t] (est_counts),
 * tpm[r*n_tx + t], iters[r]. */
int hem_solve_batch(hem_problem *p, uint32_t n_rep, const uint32_t *unique,
                    const double *cls_count, const hem_opts *o, double *est,
                    double *tpm, uint32_t *iters, hem_batch_stats *st);

/* hera's -b bootstrap (src/EM.c:123-150, 233-261) with multinomial
 * semantics: replicate r redraws all N = sum(unique) + sum(class counts)
 * fragments with replacement (Philox-4x32-10 key seeds[r]), giving new
 * unique and class counts, and solves them.  Outputs as hem_solve_batch. */
int hem_bootstrap(hem_problem *p, const uint32_t *unique, uint32_t n_rep,
                  const uint64_t *seeds, const hem_opts *o, double *est,
                  double *tpm, uint32_t *iters, hem_batch_stats *st);

/* The resampled counts hem_bootstrap would solve: unique_out[r*n_tx + t],
 * cls_out[r*n_cls + c] (classes in the caller's order). */
int hem_bootstrap_counts(hem_problem *p, const uint32_t *unique,
                         uint32_t n_rep, const uint64_t *seeds,
                         uint32_t *unique_out, double *cls_out);

/* Bootstrap with hera's own sampler (create_distribution and
 * bootstrap_sample, src/EM.c:98-150), replacing the replicate loop of
 * write_result (src/EM.c:233-255): each replicate redraws the N unique
 * reads as N - 1 draws of slot rand_r() % n for n = N-1 down to 1 (a read
 * slot belongs to its transcript through the running sums of unique);
 * class counts are not resampled.  rand_r is glibc's (three steps of the
 * LCG x -> 1103515245 x + 12345 per call); replicate r (0-based) starts
 * from the state 3 (N-1) r steps past seed, so seed 0 gives, bit for bit,
 * the counts of hera's bootstrap run on one thread (its global seed starts
 * at 0, src/hash_align.c:33, and its replicates follow each other), and
 * any seed gives replicates that do not depend on scheduling.  Then every
 * replicate is solved as by hem_bootstrap.  One-sample handles; N < 2^32;
 * n_rep <= 65535. */
int hem_bootstrap_hera(hem_problem *p, const uint32_t *unique,
                       uint32_t n_rep, uint32_t seed, const hem_opts *opts,
                       double *est, double *tpm, uint32_t *iters,
                       hem_batch_stats *st);
/* The unique counts hem_bootstrap_hera would solve: unique_out[r*n_tx + t]. */
int hem_bootstrap_hera_counts(hem_problem *p, const uint32_t *unique,
                              uint32_t n_rep, uint32_t seed,
                              uint32_t *unique_out);

/* ---- per-kernel timing --------------------------------------------------
 * Runs the same solve as hem_solve, one iteration at a time outside the
 * graph, with a HIP event after every launch on the solve's stream, and
 * reports the device time and algorithmic bytes of each kernel kind.
 * Kinds: 0 class pass (denominators), 1 transcript pass (gather + M-step),
 * 2 norm pass 1, 3 norm pass 2 (tree partial sums, or the exact sums),
 * 4 extrapolation, 5 iteration end, 6 tile maps u -> u1 -> u2 (FP64 solves
 * over component tiles), 7 tile extrapolation + map + stop test. */
#define HEM_KERNEL_KINDS 8
typedef struct {
	uint32_t iters;
	uint32_t launches[HEM_KERNEL_KINDS];
	double   ms[HEM_KERNEL_KINDS];     /* summed device time per kind      */
	double   bytes[HEM_KERNEL_KINDS];  /* algorithmic bytes per launch     */
} hem_kernel_profile;

int hem_profile_solve(hem_problem *p, const uint32_t *unique,
                      const hem_opts *o, hem_kernel_profile *out);

/* The same for a batch: the solves of hem_solve_batch (same arguments),
 * one iteration at a time outside the graph with a HIP event after every
 * launch, reporting per kernel kind the launches, device time and
 * algorithmic bytes summed over launches.  Bytes count u32 indices and
 * FP32 values: per EM map the class structure once, 8 nnz + 4 n_cls +
 * 8 n_tx, plus 12 n_cls + 16 n_tx for each replicate the launch advances;
 * an exact-sum pass 16 n_tx per replicate.  Kinds: 0 first iteration's
 * maps u -> u1 -> u2 (2 maps), 1 norm pass 1, 2 norm pass 2 (exact sums,
 * chains or tree alpha), 3 extrapolated map with the stop test and the
 * next iteration's two maps (3 maps), 4 iteration end. */
#define HEM_BATCH_KINDS 5
typedef struct {
	uint32_t iters;        /* iterations run (the slowest replicate's k)  */
	uint32_t chunk, n_tiles, sum_reruns;
	uint64_t rep_iters;    /* replicate-iterations (sum of every k)       */
	uint32_t launches[HEM_BATCH_KINDS];
	double   ms[HEM_BATCH_KINDS];     /* summed device time per kind      */
	double   bytes[HEM_BATCH_KINDS];  /* algorithmic bytes, all launches  */
} hem_batch_profile;

int hem_profile_batch(hem_problem *p, uint32_t n_rep, const uint32_t *unique,
                      const double *cls_count, const hem_opts *o,
                      hem_batch_profile *out);

/* Device-to-device copy of `bytes` repeated `reps` times on a fresh stream:
 * read + write bytes per second.  No counterpart in the reference; the
 * achievable-bandwidth reference for the kernels' figures. */
int hem_copy_bandwidth(size_t bytes, int reps, double *gb_per_s);

/* ---- exact left-to-right sums ---------------------------------------------
 * The reductions on the reference's path, bitwise, over n_rows rows of n
 * values (x[r*n + i], host memory): HEM_SUM_PLAIN out[r] = the sum
 * ((+0 + x_0) + x_1) + ... in index order (src/EM.c:362-363); HEM_SUM_NORM_SD
 * out[r] = Norm(x_r) (src/hash_align.c:150-164: mean, centred sum of
 * squares, / (n-1), sqrt); HEM_SUM_SQUARES out[r] = the same ordered sum of
 * x_i * x_i.  Computed in parallel chunks by chunk programs (see
 * HEM_NORM_EXACT), or, with HEM_SUM_WINDOWS or'ed into mode, from windows
 * of speculated starts (HEM_NORM_EXACT_WIN); reruns, if not NULL, receives
 * the number of chunks (or program records) added again value by value
 * from their exact start.  No counterpart in the reference as an entry
 * point. */
#define HEM_SUM_PLAIN   0
#define HEM_SUM_NORM_SD 1
#define HEM_SUM_SQUARES 2
#define HEM_SUM_WINDOWS 16  /* or'ed into mode: the EXACT_WIN method */
int hem_exact_sums(const double *x, uint32_t n, uint32_t n_rows, int mode,
                   double *out, uint32_t *reruns);

/* The same sums without any allocation, for callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops.  No
 * counterpart in the reference.  x (n_rows * n values), out (n_rows) and
 * reruns (one word, or NULL) are device pointers; the work is enqueued on
 * `stream` (a hipStream_t; NULL: the default stream) and the call returns
 * without waiting for it.  The scratch is a fixed array in device memory
 * (4.9 MB, held on each device the library runs on until the process
 * ends): n_rows * ceil(n / L) chunks, L = 256 for n <= 16384 and 1024
 * above, must not exceed 4096 (HEM_ERR_RANGE otherwise).  Calls on one
 * device share that array, so calls that overlap on a device race: the
 * caller orders them (one stream, or events).  Results are bitwise those
 * of hem_exact_sums (the window method, as with HEM_SUM_WINDOWS). */
int hem_exact_sums_noalloc(const double *x, uint32_t n, uint32_t n_rows,
                           int mode, double *out, uint32_t *reruns,
                           void *stream);

/* ---- abundance tables ------------------------------------------------------
 * hera's text outputs of write_result (src/EM.c:152-205), from host arrays
 * (no device work).  names: n fixed-width blocks of name_len bytes (a
 * block's name ends at its first NUL), as REF_INF->ref_name; est and tpm
 * as EM_val.overlap and .grama.  Rows are "%s\t%u\t%u\t%f\t%f\t%f\n" of
 * name, unique count, length, effective length, est_counts, tpm after the
 * header "#target_id\tunique_map\tlength\teff_length\test_counts\ttpm".
 * Replaces the abundance.tsv part of write_result (src/EM.c:176-185). */
int hem_write_abundance(const char *path, uint32_t n, const char *names,
                        uint32_t name_len, const uint32_t *unique,
                        const uint32_t *len, const double *eff_len,
                        const double *est, const double *tpm);
/* One row per run of consecutive transcripts with the same gene id
 * (gene[i] < n_gene, as GENE_MAP->gene): est_counts and tpm summed in
 * float, named by the run's last transcript's name block and the gene's
 * name block (l_gene bytes each, as GENE_MAP->gene_name), both from their
 * 17th byte; header "#gene_id\tgene_name\test_counts\ttpm".  Replaces the
 * abundance.gene.tsv part (src/EM.c:187-205); the reference's read of
 * gene[n] past the array at the last transcript is not reproduced (the
 * array's end ends the run).  name_len and l_gene must exceed 16. */
int hem_write_gene_abundance(const char *path, uint32_t n, const char *names,
                             uint32_t name_len, const uint32_t *gene,
                             const char *gene_names, uint32_t l_gene,
                             uint32_t n_gene, const double *est,
                             const double *tpm);
/* Both files under the reference's names: out_dir, '/' when it does not
 * end in one, prefix, '_' when prefix is not empty, then "abundance.tsv"
 * and (when gene != NULL) "abundance.gene.tsv" (src/EM.c:161-187).  The
 * HDF5 file of src/EM.c:208-231 is not written.  Errors: HEM_ERR_INVALID,
 * HEM_ERR_IO (file), HEM_ERR_NOMEM. */
int hem_write_result(const char *out_dir, const char *prefix, uint32_t n,
                     const char *names, uint32_t name_len,
                     const uint32_t *unique, const uint32_t *len,
                     const double *eff_len, const uint32_t *gene,
                     const char *gene_names, uint32_t l_gene,
                     uint32_t n_gene, const double *est, const double *tpm);

/* ---- hera-EM's simplified-model SQUAREM -------------------------------------
 * hera-EM's EM after its full-model loops (s_squarem, hera-EM/em.c:418-426
 * and 438-452) on a class table whose classes are groups of identical reads:
 * count n_c = the group's size, members = the reads' alignments (transcript,
 * member_weight = c_prob; NULL = 1), and the noise component as transcript
 * noise_tx, a member of every class with weight 4^-read_length.  The state
 * p is a distribution over all n_tx entries (the noise included); lengths
 * and unique counts of the table are not used.  Each loop runs three EMs of
 * the simple model (hera-EM/em.c:75-157, smodel.h:15-45): per class, total
 * = locked mass P0 + the live members' p_t w_k, member posteriors post =
 * (p_t w_k) / total; a member adds n_c post to its transcript unless post <
 * filter_threshold or its transcript is locked; the second EM of each loop
 * filters (removes such members for good, moving a locked member's p_t w_k
 * into P0; removes a class left with no contributing member, banking n_c
 * times its noise posterior, added to the noise in every later E-step).
 * M-step (transdis.c:20-30, 54-61): locked entries keep their value, the
 * others are rescaled to sum to 1 minus the locked mass.  Between the
 * second and third EM the unlocked entries are extrapolated (L2 step
 * gamma^2 = |r|^2 / |v|^2, values <= 1e-300 replaced by u2 / 10,
 * utility.c:64-107), then an unlocked entry locks when it is < 1e-300 or
 * moved by less than lock_threshold times itself (transdis.c:97-115).  The
 * loop ends when no entry with p3 >= stop_min_value changed by at least
 * stop_rel_change relative to p3 (is_stop, transdis.c:123-130) or at
 * max_iter (hera-EM has no cap).  Sums run in index / class order on the
 * device (hera-EM's per-thread joins have no fixed order).  p0: start
 * (NULL: uniform 1 / n_tx, trans_uniform_init).  One-sample handles, FP64. */
typedef struct {
	uint32_t noise_tx;          /* noise component, or >= n_tx for none     */
	uint32_t max_iter;          /* 1000                                      */
	int      locking;           /* 1 (hera-EM)                               */
	int      filtering;         /* 1 (hera-EM)                               */
	double   lock_threshold;    /* 1e-5, LOCK_THRESHOLD (utility.h:13)       */
	double   filter_threshold;  /* 1e-6, FILTERING_THRESHOLD (utility.h:14)  */
	double   stop_rel_change;   /* 1e-3, THRESHOLD1 (utility.h:10)           */
	double   stop_min_value;    /* 1e-7, THRESHOLD2 (utility.h:11)           */
} hem_hera_em_opts;

typedef struct {
	uint32_t iters;             /* SQUAREM loops (3 EMs each)                */
	uint32_t converged;
	uint32_t n_locked;          /* entries locked at the end                 */
	uint32_t removed_members;   /* members filtered out                      */
	uint32_t removed_classes;   /* classes removed (noise banked)            */
	double   ms;                /* device time of the loop                   */
} hem_hera_em_stats;

/* hera-EM's constants; noise_tx = 0xFFFFFFFF (none). */
void hem_hera_em_opts_default(hem_hera_em_opts *o);
int hem_solve_hera_em(hem_problem *p, const double *p0,
                      const hem_hera_em_opts *o, double *p_out,
                      hem_hera_em_stats *st);

/* ---- adjoint of the log-likelihood ----------------------------------------
 * No counterpart in the reference, which has no derivatives; the objective
 * is the one whose EM map is src/EM.c:271-302.  With pi = softmax(theta),
 * w_t = exp(-log_efflen_t), x_t = pi_t w_t, the table's class counts n_c and
 * the unique counts u_t as singleton classes:
 *   L = sum_c n_c log sum_{t in c} x_t + sum_t u_t log x_t
 *   dL/dtheta_t = E_t - N pi_t,   dL/dlog_efflen_t = -E_t,
 * where E_t = x_t sum_{c ni t} n_c / D_c + u_t are the expected counts of
 * the E-step and N = sum_t E_t.  At a fixed point of hera's EM,
 * theta = log(lambda * len) (up to a constant) and dL/dtheta = 0.
 * precision HEM_PREC_FP64, or HEM_PREC_FP32 for x and D_c in FP32 with
 * every sum of the adjoint in FP64.  Outputs may be NULL. */
typedef struct {
	double loglik;   /* L                                                 */
	double n_frag;   /* N                                                 */
	double ms;       /* device time (uploads of the point included)       */
	double ms_adjoint; /* device time of the class and transcript passes
	                      (one tile launch, or the two global passes)     */
} hem_grad_stats;

int hem_grad(hem_problem *p, const uint32_t *unique, const double *theta,
             const double *log_efflen, int precision, double *g_theta,
             double *g_log_efflen, hem_grad_stats *st);

/* The gradient of every sample of a multi-sample handle (hem_samples_create;
 * a one-sample handle works too) in one pass: sample q's objective L_q with
 * pi = softmax(theta) within the sample, its gradient E_t - N_q pi_t and
 * -E_t in the sample's columns of g_theta and g_log_efflen, loglik[q] and
 * n_frag[q] = N_q (n_samples entries each).  Each sample's values are what
 * hem_grad gives on the sample's own handle, up to rounding (the softmax
 * and the sums run in another order).  Needs every component within the
 * tile limits (HEM_ERR_RANGE otherwise); st sums over samples.  Outputs
 * may be NULL. */
int hem_grad_samples(hem_problem *p, const uint32_t *unique,
                     const double *theta, const double *log_efflen,
                     int precision, double *g_theta, double *g_log_efflen,
                     double *loglik, double *n_frag, hem_grad_stats *st);

/* Hessian-vector product in theta, FP64: hv = (d^2 L / d theta^2) v. */
int hem_hvp(hem_problem *p, const uint32_t *unique, const double *theta,
            const double *log_efflen, const double *v, double *hv);

/* Standard errors of the top_k abundances pi_j (largest pi first, ties by
 * id) by the delta method with the observed information F = -H:
 * se_j = sqrt(b_j . x_j), F x_j = b_j = pi_j (e_j - pi), solved for all
 * top_k right-hand sides at once by cg_steps of Jacobi-preconditioned
 * conjugate gradients over the component tiles (damping adds damping * I
 * to F).  Outputs: idx[top_k] transcript ids, se[top_k], resid[top_k] =
 * |r| / |b| after the last step.  Needs the component tiling
 * (hem_problem_tiles n_tiles > 0). */
typedef struct {
	uint32_t top_k, cg_steps, chunk;
	double   resid_max;          /* largest relative residual             */
	double   ms_setup;           /* gradient, constants, selection        */
	double   ms_cg;              /* device time of the CG steps           */
	double   ms_total;           /* host wall time of the call            */
} hem_se_stats;

int hem_stderr(hem_problem *p, const uint32_t *unique, const double *theta,
               const double *log_efflen, uint32_t top_k, uint32_t cg_steps,
               double damping, uint32_t *idx, double *se, double *resid,
               hem_se_stats *st);

#ifdef __cplusplus
}
#endif
#endif /* HERA_EM_H_ */
