/*
 * hem_dev.h -- the thin C ABI between the C host code and the HIP side.
 *
 * Everything the host needs from the device goes through these functions:
 * plain pointers, sizes and int codes (0 = success, otherwise a HIP error
 * number; hemd_error_string() names it).  The host C files never include a
 * HIP header.
 */
#ifndef HEM_DEV_H_
#define HEM_DEV_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ------------------------------------------------------------ */
int hemd_init(int *device, char *arch, size_t arch_len, int *n_cu);
const char *hemd_error_string(int code);
int hemd_malloc(void **p, size_t n);
int hemd_free(void *p);
int hemd_host_alloc(void **p, size_t n);   /* pinned */
int hemd_host_free(void *p);
int hemd_stream_create(void **s);
int hemd_stream_destroy(void *s);
int hemd_stream_sync(void *s);
int hemd_h2d(void *dst, const void *src, size_t n, void *s);
int hemd_d2h(void *dst, const void *src, size_t n, void *s);
int hemd_d2d(void *dst, const void *src, size_t n, void *s);
int hemd_event_create(void **e);
int hemd_event_destroy(void *e);
int hemd_event_record(void *e, void *s);
int hemd_event_sync(void *e);
int hemd_event_elapsed(void *e0, void *e1, float *ms);
int hemd_graph_begin(void *s);
int hemd_graph_end(void *s, void **exec);
int hemd_graph_launch(void *exec, void *s);
int hemd_graph_destroy(void *exec);
int hemd_last_launch_error(void);

/* ---- solver control block (device memory) -------------------------------
 * One per EM instance.  Every iteration kernel reads `done` first and does
 * nothing once it is set, so a captured chunk of iterations can run past
 * convergence without changing the result. */
typedef struct {
	uint32_t k;           /* SQUAREM iterations completed                  */
	uint32_t done;        /* loop finished                                 */
	uint32_t maps;        /* EM maps applied                               */
	uint32_t converged;   /* stop rule fired                               */
	uint32_t min_iter, max_iter;
	uint32_t fallbacks;
	uint32_t reruns;      /* chunks of the exact sums run again (misses)   */
	uint32_t stopbad;     /* set (OR) by the third map when some
	                         transcript fails the stop test; read and
	                         cleared by the iteration end              */
	uint32_t stopbad2;    /* monotonicity fallback: the same for the
	                         second map (u2 against u1)                */
	uint32_t fb;          /* monotonicity fallback: this iteration keeps
	                         u2                                        */
	double   alpha;       /* last step length                              */
	double   stop_min, stop_rel;
	double   mean_r, mean_v;  /* first pass of the sequential norms        */
	double   lam_sum;         /* sequential sum for the finalisation       */
	uint32_t *hdone;          /* pinned host word set to 1 with done (the
	                             host's loop reads it without a copy), or
	                             NULL                                  */
	uint32_t *hk;             /* pinned host word: k after each iteration
	                             end, or NULL                          */
} hemk_ctl;

/* Device view of one problem (all pointers are device pointers). */
typedef struct {
	uint32_t n_tx, n_cls;
	uint64_t nnz;
	const uint32_t *cls_off;   /* n_cls+1, CSR over classes                */
	const uint32_t *cls_ref;   /* nnz member transcript ids                */
	const float    *cnt32;     /* n_cls class counts (FP32 copy)           */
	const double   *cnt64;     /* n_cls class counts                       */
	const uint32_t *tx_off;    /* n_tx+1, CSC over transcripts             */
	const uint32_t *tx_ref;    /* nnz class ids, ascending per transcript  */
	const double *cls_w;       /* nnz member weights (CSR), or NULL        */
	const double *tx_w;        /* the same weights in CSC order, or NULL   */
	const uint32_t *len;       /* n_tx transcript lengths                  */
} hemk_prob;

/* Step options fixed for one solve. */
typedef struct {
	int step;        /* 0 sample sd, 1 L2                                 */
	int clamp;       /* 0 zero, 1 u2/10                                   */
	int alpha_clamp; /* alpha <= -1                                       */
	int alpha_given; /* 1: alpha already in ctl (sequential norms); 2:
	                    from the two sums in part[0], part[1] (exact
	                    sums by speculation)                           */
} hemk_step;

#define HEMK_NORM_BLOCKS 512

/* ---- component tiles (device view) ---------------------------------------
 * See host/tiles.c.  desc holds {tx_base, n_tx, cls_base, n_cls, member
 * base, members, class-list base, record offset} per tile (HEMK_TILE_DESC
 * words). */
#define HEMK_TILE_DESC 8
typedef struct {
	uint32_t n_tiles, n_tx, n_cls;
	uint32_t max_tx, max_cls, max_nnz;
	const uint32_t *desc;
	const uint32_t *tx;     /* tile position -> transcript id            */
	const uint32_t *len;    /* tile position -> length                   */
	const uint32_t *coff;   /* class position -> member offset (+1)      */
	const uint32_t *kmem;   /* member -> (local class << 8) | local tx    */
	const uint32_t *toff;   /* tile position -> class-list offset (+1)   */
	const uint16_t *tcls;   /* class-list entry -> local class           */
	const uint32_t *txq;    /* transcript id -> tile position            */
	const uint32_t *rec;    /* per-tile records (host/tiles.c), 16-byte
	                           units; tile i's at desc[8i + 7]          */
	uint32_t max_rec;       /* bytes of the largest record               */
	uint32_t weighted;      /* records carry member weights             */
} hemk_tiles;

/* Per-instance control of a batched solve.  A handle holds ns samples
 * (independent class tables side by side: sample s owns transcript ids and
 * tile positions [tx0_s, tx0_s + n_tx_s) and tiles [tile0_s, tile0_s +
 * n_tiles_s)); each runs R replicates, and instance i = s * R + r has its
 * own EM (control arrays of length ns * R, indexed by i). */
typedef struct {
	uint32_t R;
	uint32_t min_iter, max_iter;
	uint32_t chunk;     /* replicates per tile workgroup (RC)           */
	double   stop_min, stop_rel;
	uint32_t *k, *done, *conv, *notconv;
	double   *alpha, *mean_r, *mean_v, *lam_sum;
	double   *q_r, *q_v;  /* second-pass sums of the sequential norms */
	uint32_t *active;   /* [ns]: replicates of each sample still
	                       iterating */
	uint32_t *alist;    /* per sample, R + 1 words: count, then the
	                       replicates still iterating, ascending        */
	uint32_t ns;              /* samples (1: a single class table)    */
	const uint32_t *tsamp;    /* tile -> sample (NULL when ns == 1)   */
	const uint32_t *txsamp;   /* transcript id or tile position ->
	                             sample (NULL when ns == 1)           */
	const uint32_t *sinfo;    /* per sample {tx0, n_tx, tile0, n_tiles}
	                             (NULL when ns == 1)                  */
} hemk_rctl;

/* Largest replicate chunk (lanes per tile block) that fits the LDS budget
 * for the given tiling and count width (4 or 8 bytes). */
int hemk_tile_chunk(const hemk_tiles *t, uint32_t R, int cnt_bytes);

/* Tile map kernels.  Layouts, rows in tile position order q (transcripts
 * of a tile are consecutive; txq maps a transcript id to q): state A/B/C
 * [q][R] FP64 (replicate fastest), unique counts [q][R] u32, class counts
 * [class tile position][R] u32.  f32: the FP32 bulk form of every map
 * (FP32 class sums, class factors n_c / D_c and terms, FP64 transcript
 * sums, as hemk_ptile_map12 computes it).
 *   map12: u1 = M(A) -> B, u2 = M(B) -> C; with part != NULL also the
 *          per-(tile, replicate) sums of r and v (Chan form) for the
 *          parallel norms, part[(r*n_tiles + tile)*4 ...]; with xr the
 *          r and v values in replicate-major rows xr/xv[r*T + t], t the
 *          transcript id, for the sequential norms.
 *   ex3:   extrapolate from A, B, C (alpha[r] when st.alpha_given, else
 *          from q_r/q_v), map once -> A, stop test -> notconv[r]. */
int hemk_tile_map12(const hemk_tiles *t, hemk_rctl rc, int chunk,
                    const uint32_t *cnt, int f32, const uint32_t *uniq,
                    const double *A, double *B, double *C, double *part,
                    double *xr, double *xv, uint32_t T, int step, void *s);
/* ex3 with next != 0 also maps twice more for the next iteration:
 * B = M(A), C = M(B) and the r, v rows (xr, xv) or tree partials (part),
 * as map12 does. */
int hemk_tile_ex3(const hemk_tiles *t, hemk_rctl rc, int chunk,
                  const uint32_t *cnt, int f32, const uint32_t *uniq,
                  double *A, double *B, double *C, uint32_t T,
                  hemk_step st, int next, double *part, double *xr,
                  double *xv, void *s);
/* Sequential sums of a batch over replicate-major rows, one left-to-right
 * chain per (instance, row) block over the instance's sample's transcripts
 * (row r, ids tx0_s ..), in the reference's order over t:
 *   phase 0: lam_sum[i] = sum of xr row (every instance)
 *   phase 1: mean_r/mean_v = sums / n_tx_s
 *   phase 2: q_r/q_v = centred sums of squares    phase 3: sums of squares
 * Phases 1-3 skip finished instances. */
int hemk_batch_chain(hemk_rctl rc, uint32_t T, const double *xr,
                     const double *xv, int phase, void *s);
/* Parallel norms of a batch: fold the tile partials of each replicate in
 * tile order (Chan et al. for the sd form) -> alpha[r]. */
int hemk_batch_alpha(const hemk_tiles *t, hemk_rctl rc, const double *part,
                     uint32_t T, hemk_step st, void *s);
int hemk_batch_iter_end(hemk_rctl rc, void *s);
int hemk_batch_reset(hemk_rctl rc, uint32_t T, double *A, void *s);
/* X[r*T + t] = A[txq[t]*R + r]: the state as one row per replicate in
 * transcript order. */
int hemk_batch_rows(hemk_rctl rc, uint32_t T, const uint32_t *txq,
                    const double *A, double *X, void *s);
/* est[r][t] = X[r][t] len[t]; tpm[r][t] = X[r][t] 1e6 / lam_sum[r] (zero
 * when the sum is zero); lam_sum from hemk_batch_chain phase 0. */
int hemk_batch_out(hemk_rctl rc, uint32_t T, const double *X,
                   const uint32_t *len, double *est, double *tpm, void *s);
/* out = transpose of in (rows x cols, row-major) for u32 count blocks. */
int hemk_transpose_u32(const uint32_t *in, uint32_t rows, uint32_t cols,
                       uint32_t *out, void *s);

/* ---- bootstrap resampling (resample.hip) --------------------------------
 * Categories k < K with integer weights w[k], N = sum w < 2^32, split into B
 * chunks [chunk_start[j], chunk_start[j+1]) with weight prefix chunk_cw
 * (chunk_cw[B] = N; a chunk holds at most HEMK_RS_CHUNK categories).  Draws
 * Multinomial(N, w/N) for every replicate r (Philox key seeds[r]) into
 * out[r*K + k].  m: R*B u32 scratch. */
#define HEMK_RS_CHUNK 2048
#define HEMK_RS_MAXB 7000
int hemk_resample(const uint32_t *w, uint32_t K, const uint32_t *chunk_start,
                  const uint32_t *chunk_cw, uint32_t B, uint32_t N,
                  const uint64_t *seeds, uint32_t R, uint32_t *m,
                  uint32_t *out, void *s);
/* hera's sampler (src/EM.c:98-150): rows out[R][T + C] get the unique
 * counts of N - 1 draws each (cum: running sums of the unique counts, N =
 * cum[T-1]) and the class counts cls; replicate r's rand_r state starts
 * 3 (N-1) r LCG steps past seed.  bt: ((N-1) >> sh) + 1 words of
 * scratch. */
int hemk_hera_resample(const uint32_t *cum, uint32_t T, uint32_t N,
                       uint32_t *bt, uint32_t sh, const uint32_t *cls,
                       uint32_t C, uint32_t seed, uint32_t R, uint32_t *out,
                       void *s);
int hemk_fill_u32(uint32_t *p, size_t n, uint32_t v, void *s);
/* One sample's resampled rows in[R][Tq + Cq] into out[R][K]: unique counts
 * at off_u.., class counts at off_c.. */
int hemk_scatter_counts(const uint32_t *in, uint32_t R, uint32_t Tq,
                        uint32_t Cq, uint32_t *out, uint32_t K,
                        uint32_t off_u, uint32_t off_c, void *s);
/* out[q*R + r] = in[idx[q]*R + r] */
int hemk_gather_rows_u32(const uint32_t *in, const uint32_t *idx,
                         uint32_t rows, uint32_t R, uint32_t *out, void *s);

/* ---- adjoint of the log-likelihood (grad_kernels.hip) --------------------
 * part: 3 * HEMK_NORM_BLOCKS doubles; scal[0] = L, scal[1] = N. */
int hemk_grad_softmax(const double *theta, const double *ell, uint32_t T,
                      double *pi, double *w, double *x, float *x32,
                      double *part, void *s);
int hemk_grad_passes(const hemk_prob *p, int fp32, const double *x,
                     const float *x32, const uint32_t *unique,
                     const double *pi, double *rho, double *D, double *kappa,
                     double *E, double *fd, double *gth, double *gell,
                     double *a, double *part, double *scal, void *s);
/* hv = H v (FP64); kd: n_cls doubles; part: HEMK_NORM_BLOCKS doubles. */
int hemk_hvp(const hemk_prob *p, const double *x, const double *a,
             const double *pi, const double *kappa, const double *scal,
             const double *v, double *kd, double *part, double *hv, void *s);
/* The gradient's last pass (g_theta, g_ell, a, the Fisher diagonal and
 * scal) from np partials each of sum n_c log D_c (part[0..np)), N
 * (part[np..2np)) and sum u_t log x_t (part[2np..3np)); part[3np..3np+3)
 * is scratch for their totals. */
int hemk_grad_fin(const double *E, const double *pi, const uint32_t *unique,
                  uint32_t T, double *part, uint32_t np, double *gth,
                  double *gell, double *a, double *fd, double *scal, void *s);
/* The adjoint's class and transcript passes over the single-problem tiles
 * (the classes of a tile in LDS): kappa_p by class tile position, E and fd
 * by transcript id, as hemk_grad_passes computes them, and per-tile
 * partials for hemk_grad_fin (np = n_tiles).  cntq: class counts by tile
 * position; x (or x32 with fp32), unique: by transcript id. */
int hemk_ptile_grad(const hemk_tiles *t, const double *cntq, int fp32,
                    const double *x, const float *x32,
                    const uint32_t *unique, double *kappa_p, double *E,
                    double *fd, double *part, void *s);
/* Several samples side by side (info: per sample {first transcript,
 * transcripts, first tile, tiles} of the single-problem tiling): pi, w, x
 * with a softmax within each sample; then, from the tile passes' partials,
 * the per-sample totals tot[3q..3q+2] = {sum n_c log D_c, N_q,
 * sum u_t log x_t} and the gradient with each transcript's sample's N
 * (txsamp: transcript -> sample, NULL for one sample). */
int hemk_grad_softmax_seg(const double *theta, const double *ell,
                          const uint32_t *info, uint32_t ns, double *pi,
                          double *w, double *x, float *x32, void *s);
int hemk_grad_fin_seg(const double *E, const double *pi,
                      const uint32_t *unique, uint32_t T,
                      const uint32_t *txsamp, const uint32_t *info,
                      uint32_t ns, const double *part, uint32_t np,
                      double *tot, double *gth, double *gell, double *a,
                      double *fd, void *s);
/* HEMK_NORM_BLOCKS block partials of a . b (T values). */
int hemk_dot_parts(const double *a, const double *b, uint32_t T,
                   double *part, void *s);
/* hv = H v over the tiles: kappa_p by class tile position, pv_part holds
 * HEMK_NORM_BLOCKS partials of pi . v (hemk_dot_parts), scal[1] = N. */
int hemk_ptile_hvp(const hemk_tiles *t, const double *x, const double *v,
                   const double *kappa_p, const double *a, const double *pi,
                   const double *pv_part, const double *scal, double *hv,
                   void *s);

/* ---- standard errors by batched CG (se_kernels.hip) ----------------------
 * Vectors [q][K] in tile position order; per-direction partials [rb][K]
 * over hemk_se_rowblocks(T) row blocks, folded by hemk_cg_fold. */
int hemk_se_chunk(const hemk_tiles *t, uint32_t K);
uint32_t hemk_se_rowblocks(uint32_t T);
/* Q = F P = -(H P) + delta P for every direction (pv: pi . P per
 * direction, scal[1] = N), and per (tile, direction) partials of P . F P
 * in part[tile * K + k], folded over n_tiles by hemk_cg_fold. */
int hemk_se_tile(const hemk_tiles *t, int chunk, uint32_t K, const double *xq,
                 const double *aq, const double *kq, const double *piq,
                 const double *pv, const double *scal, double delta,
                 const double *P, double *Q, double *part, void *s);
int hemk_cg_init(const double *piq, const double *mq, const uint32_t *jq,
                 const double *pj, uint32_t T, uint32_t K, double *X,
                 double *R, double *P, double *prz, double *prr, double *ppv,
                 void *s);
int hemk_cg_xr(const double *mq, const double *rz, const double *pfp,
               const double *rr, const double *bb, uint32_t T, uint32_t K,
               const double *P, const double *FP, double *X, double *R,
               double *prz, double *prr, void *s);
int hemk_cg_p(const double *piq, const double *mq, const double *rz_old,
              const double *rz_new, uint32_t T, uint32_t K, const double *R,
              double *P, double *ppv, void *s);
int hemk_cg_pix(const double *piq, uint32_t T, uint32_t K, const double *X,
                double *ppx, void *s);
int hemk_cg_fold(const double *part, uint32_t nrb, uint32_t K, double *out,
                 void *s);
int hemk_cg_out(const double *X, const uint32_t *jq, const double *pj,
                const double *pix, const double *rr, const double *bb,
                uint32_t K, double *se, double *resid, void *s);
int hemk_gather_f64(const double *in, const uint32_t *idx, uint32_t n,
                    double *out, void *s);
/* out[idx[i]] = in[i] (idx a permutation of its range) */
int hemk_scatter_f64(const double *in, const uint32_t *idx, uint32_t n,
                     double *out, void *s);
int hemk_precond(const double *fdq, const double *piq, const double *scal,
                 double delta, uint32_t n, double *mq, void *s);

/* ---- kernel launchers --------------------------------------------------- */
/* lam_init: fill lam64 (and lam32 if non-NULL) with `value`, or copy from
 * init64 when non-NULL. */
int hemk_init_state(uint32_t n, double *lam64, float *lam32, double value,
                    const double *init64, void *s);
/* ctl_reset: hdone and hk (pinned host memory, or NULL) receive 1 when the
 * loop ends and k after each iteration end; the caller clears them
 * before. */
int hemk_ctl_reset(hemk_ctl *ctl, uint32_t min_iter, uint32_t max_iter,
                   double stop_min, double stop_rel, uint32_t *hdone,
                   uint32_t *hk, void *s);

/* CSR pass: rho_c = n_c / sum_{t in c} lam_t (0 if the sum is 0).
 * prec 0: lam32/rho32; prec 1: lam64/rho64.  group = lanes per class. */
int hemk_denom(const hemk_prob *p, int prec, const void *lam, void *rho,
               int group, const hemk_ctl *ctl, void *s);

/* CSC pass + M-step: s_t = sum_{c ni t} rho_c, out_t = (u_t + lam_t s_t) /
 * len_t; FP32 shadow written when out32 != NULL; when stop_flag != NULL a
 * block with a transcript that fails the stop test ORs 1 into it. */
int hemk_gather_mstep(const hemk_prob *p, int prec, const void *rho,
                      const uint32_t *unique, const double *lam64,
                      double *out64, float *out32, uint32_t *stop_flag,
                      int group, const hemk_ctl *ctl, void *s);
uint32_t hemk_gather_blocks(const hemk_prob *p, int group);

/* SQUAREM: norms of r = u1-u, v = (u2-u1)-r, step length, extrapolation. */
int hemk_norm_pass1(uint32_t n, const double *u, const double *u1,
                    const double *u2, double *part1, int step,
                    const hemk_ctl *ctl, void *s);
int hemk_norm_pass2(uint32_t n, const double *u, const double *u1,
                    const double *u2, const double *part1, double *part2,
                    int step, const hemk_ctl *ctl, void *s);
int hemk_extrapolate(uint32_t n, const double *u, const double *u1,
                     const double *u2, const double *part2, double *out64,
                     float *out32, hemk_step st, hemk_ctl *ctl, void *s);

/* Sequential norms, bitwise the reference's order: one wave lane runs
 * each left-to-right sum of src/hash_align.c:154-160 while other waves
 * stage the next chunk in LDS.
 *   phase 1 (sd): mean_r = (sum r)/n, mean_v = (sum v)/n into ctl
 *   phase 2 (sd): centred sums -> n_r, n_v -> alpha into ctl
 *   phase 3 (L2): sum r^2, sum v^2 -> alpha into ctl
 *   phase 0:      lam_sum = sum u (the finalisation sum, src/EM.c:362-363),
 *                 runs even when ctl->done is set. */
int hemk_seq_norm(uint32_t n, const double *u, const double *u1,
                  const double *u2, int phase, hemk_step st, hemk_ctl *ctl,
                  void *s);

/* ---- one EM instance over component tiles (ptile_kernels.hip) -----------
 * State A, B, C in tile order; uq (unique counts) and cntq (class counts,
 * class tile positions) likewise.  map12: B = M(A), C = M(B), and r, v
 * rows in transcript order into xr, xv when xr != NULL.  ex3_map12:
 * extrapolate (alpha: st.alpha_given 1 ctl->alpha, 2 from part[0..1]),
 * A = M(.) with the stop test into ctl->stopbad, then B = M(A), C = M(B)
 * and the r, v rows as map12 does.  alpha: tree partials (nparts pairs) ->
 * ctl->alpha.  rows: out[t] = A[txq[t]].  lds: dynamic LDS of one tile. */
size_t hemk_ptile_lds(const hemk_tiles *t);
/* f32: the FP32 bulk form (class sums, class factors n_c / D_c and terms
 * in FP32; per-transcript sums, M-step, state and stop test in FP64);
 * unweighted tables only. */
/* tpart / part (map12) != NULL: the tree norms' per-tile partials of the
 * r, v just computed, [tile][4] (step form `step`), folded by
 * hemk_ptile_tree_alpha into ctl->alpha. */
int hemk_ptile_map12(const hemk_tiles *t, const double *A, double *B,
                     double *C, const uint32_t *uq, const double *cntq,
                     double *xr, double *xv, double *part, int step,
                     int f32, const hemk_ctl *ctl, void *s);
int hemk_ptile_ex3_map12(const hemk_tiles *t, double *A, double *B,
                         double *C, const uint32_t *uq, const double *cntq,
                         const double *part, uint32_t T, hemk_step st,
                         double *xr, double *xv, double *tpart, int f32,
                         hemk_ctl *ctl, void *s);
/* end != 0: first end the iteration (as hemk_iter_end), folding only if the
 * loop goes on. */
int hemk_ptile_tree_alpha(const hemk_tiles *t, const double *part,
                          uint32_t T, hemk_step st, int end, hemk_ctl *ctl,
                          void *s);
int hemk_ptile_alpha(const double *part, uint32_t nparts, uint32_t T,
                     hemk_step st, hemk_ctl *ctl, void *s);
/* Monotonicity fallback (no counterpart in the reference).  map12_fb: as
 * map12, and the second map's stop test into ctl->stopbad2.  ex3_fb: from
 * A, B, C extrapolate, lam = M(u') -> A with the stop test (ctl->stopbad),
 * and per tile the terms of ell(lam) - ell(u2) (u2 = C) as
 * log1p-differences: dpart[2 tile] = the tile's difference, dpart[2 tile
 * + 1] = the sum of its terms' magnitudes.  fb_decide: totals over tiles;
 * when the difference is below -tol times the magnitudes the iteration
 * keeps u2 (ctl->fb, ctl->fallbacks, stop test of the second map).
 * fb_apply: A = C when ctl->fb. */
int hemk_ptile_map12_fb(const hemk_tiles *t, const double *A, double *B,
                        double *C, const uint32_t *uq, const double *cntq,
                        double *xr, double *xv, double *part, int step,
                        hemk_ctl *ctl, void *s);
int hemk_ptile_ex3_fb(const hemk_tiles *t, double *A, const double *B,
                      const double *C, const uint32_t *uq,
                      const double *cntq, const double *part, uint32_t T,
                      hemk_step st, double *dpart, hemk_ctl *ctl, void *s);
int hemk_fb_decide(const double *dpart, uint32_t n_tiles, double tol,
                   hemk_ctl *ctl, void *s);
int hemk_fb_apply(uint32_t T, const double *C, double *A,
                  const hemk_ctl *ctl, void *s);
int hemk_ptile_rows(uint32_t T, const uint32_t *txq, const double *A,
                    double *out, void *s);

/* ---- hera-EM's simplified model (hera_em_kernels.hip) ---------------------
 * State of one solve on the device: the table (CSR, CSC with the member
 * index of every occurrence), the noise transcript z (>= T: none), the
 * filtering threshold, and the persistent lock / live / removed flags,
 * locked masses P0 and noise banks; rows and sums are scratch (sums:
 * B, unlocked sum, locked sum, |r|^2, |v|^2; counts: removed members,
 * removed classes, failed stop tests). */
typedef struct {
	uint32_t T, C, z;
	double thr;
	const uint32_t *off, *ref, *tx_off, *tx_ref, *occ2mem;
	const double *cnt, *w, *tx_w;
	uint8_t *lock, *alive, *dead;
	double *P0, *bank, *total;
	double *brow, *zrow, *urow, *lrow;
	double *sums;
	uint32_t *counts;
} hemk_he;
int hemk_he_occ2mem(uint32_t C, const uint32_t *off, const uint32_t *ref,
                    const uint32_t *tx_off, const uint32_t *tx_ref,
                    uint32_t *occ2mem, void *s);
int hemk_he_fill(double *x, uint32_t n, double v, void *s);
int hemk_he_rowsum(const double *x, uint32_t n, double *out, void *s);
/* E-step p -> acc (filt: the second EM of a loop, which filters) */
int hemk_he_estep(const hemk_he *h, const double *p, int filt, double *acc,
                  void *s);
/* M-step: out = locked ? p : acc (1 / sum unlocked acc) (1 - sum locked p) */
int hemk_he_mstep(const hemk_he *h, const double *p, const double *acc,
                  double *out, void *s);
/* p3 = the locked extrapolation from p1, p2, p3; then the locking rule */
int hemk_he_accelerate(const hemk_he *h, const double *p1, const double *p2,
                       double *p3, int locking, double lock_thr, void *s);
/* counts[2] += entries failing the stop test between p3 and its image p1 */
int hemk_he_stop(const hemk_he *h, const double *p3, const double *p1,
                 double stop_min, double stop_rel, void *s);

/* ---- exact left-to-right sums in parallel (spec_sum.hip) -----------------
 * Chains of n FP64 values; each chain's result is bitwise the sequential
 * sum ((+0 + y_0) + y_1) + ... of the reference's loops, computed by chunks
 * of L values: guesses of every chunk's start, W candidate starts per
 * chunk run in parallel, then an in-order lookup (a start outside its
 * window reruns that chunk).
 * Chain slot q (grid height nchain) sums row `row` of x (rows ld apart):
 *   pairs == 0: row = rep(q);   pairs == 1: row = (q & 1) * R + rep(q / 2)
 *   rep(i) = alist ? alist[1 + i] (slots with i >= alist[0] do nothing)
 *            : i.
 * y = x (IDENT), (x - mres[row] / n)^2 (CENTRED: the second pass of the
 * sample sd), or x * x (SQUARE).  The sum goes to res[row]; reruns are
 * added to *miss when miss != NULL.  With ctl != NULL every kernel returns
 * at once when ctl->done is set.  Scratch, per chain slot: csum, drift,
 * base [nchain][P]; tab [nchain][P][W] doubles and nx [nchain][P][W] bytes
 * (W at most HEMK_SPEC_W). */
#define HEMK_SPEC_W 128
#define HEMK_SPEC_LMAX 1024   /* longest chunk the kernels accept */
#define HEMK_SPEC_IDENT 0
#define HEMK_SPEC_CENTRED 1
#define HEMK_SPEC_SQUARE 2
#define HEMK_SPEC_NOLINK 0xFFu   /* nx entry: result outside next window */
typedef struct {
	uint32_t n, L, P;      /* values per chain, chunk length, chunks      */
	uint32_t W;            /* candidates per chunk: 8, 16, 32, 64 or 128  */
	uint32_t nchain;       /* chain slots (grid height)                   */
	uint32_t R;            /* pairs: row offset of the second row block   */
	int pairs, xf;
	const double *x;       /* rows; written by the pass-1 producer        */
	size_t ld;
	const uint32_t *alist;
	const double *mres;    /* CENTRED: pass-1 sums by row                 */
	double *csq;           /* IDENT, chain slot = row (any chains with chunk
	                          programs, by chain slot): when set, the chunk
	                          sums step also writes each chunk's sum of
	                          x * x here ([row][P])                    */
	const double *gsum, *gsq; /* CENTRED, chain slot = row (any chains
	                          with chunk programs, by chain slot, pass 1
	                          on the same slots): pass-1 chunk
	                          sums of x and x * x ([row][P]); when set,
	                          the drifts take each chunk's sum of
	                          (x - m)^2 as gsq - 2 m gsum + L m^2 (it only
	                          steers the guesses) and the chunk sums step
	                          is not needed                             */
	double *csum, *drift, *tab;
	long long *base;
	uint8_t *nx;           /* [nchain][P][W] links to the next window     */
	double *res;           /* sums by row                                 */
	uint32_t *miss;
	/* Chunk programs (walk_sum.hip).  When wa is set, the finishing step
	 * describes each chunk by at most L/2 records and walks them in order
	 * (no drifts, windows or links; W, tab, nx and base unused):
	 *   wa  [nchain][P][L/2] what the record adds;
	 *   wlh [nchain][P][L/2] pairs {lo, hi}: the starts it is valid for;
	 *   wrg [nchain][P][L/2] its first value's index in the chunk (bits
	 *       0-10; its values run to the next record's, or to the chunk's
	 *       end) and, in bits 11-21, 0 or the binade code plus one of the
	 *       unit 2^(code - 1075) whose multiple the start must be;
	 *   wcnt [nchain][P] records per chunk, or HEMK_WS_RAW (the chunk is
	 *       added value by value). */
	double *wa, *wlh;
	uint32_t *wrg, *wcnt;
	/* chunk programs without a seeded guess: when set ([nchain][P]
	 * words, zeroed when allocated and written by nothing else), each
	 * program publishes its chunk's sum here as one word, the double with
	 * its low 16 significand bits replaced by the launch's epoch (its own
	 * word's last epoch plus one, mod 2^16), and takes the earlier chunks'
	 * sums from their words once they carry that epoch, so no chunk sums
	 * step runs (csum is not written); every launch must cover all P
	 * chunks of a chain.  The words read as doubles are the chunk sums to
	 * a relative 2^-36, for the seeds of a centred pass */
	unsigned long long *cpub;
	/* chunk programs: when set, the walk's workgroup 0 ends the previous
	 * SQUAREM iteration on this control block (as hemk_iter_end) once its
	 * sum is written */
	hemk_ctl *iend;
} hemk_spec;
#define HEMK_WS_RAW 0xFFFFFFFFu
/* Walk-step records of launches of `chains` chains of P chunks of L. */
size_t hemk_ws_records(uint32_t L, uint32_t chains, uint32_t P);
/* The finishing step by chunk programs (called by hemk_spec_sum and
 * hemk_spec_phases when sp->wa is set). */
int hemk_ws_finish(const hemk_spec *sp, const hemk_ctl *ctl, void *s);
/* The steps of hemk_spec_sum by themselves, in this order: chunk sums,
 * drifts, then windows and resolve (produce == 0). */
#define HEMK_SPEC_PH_CSUM   1
#define HEMK_SPEC_PH_DRIFT  2
#define HEMK_SPEC_PH_FINISH 4
int hemk_spec_phases(const hemk_spec *sp, int phases, hemk_ctl *ctl,
                     void *s);
/* Chunk length for chains of n values, and the window width for a
 * transform at that length when a launch holds `chunks` chunks (chains x
 * chunks per chain). */
uint32_t hemk_spec_chunk(uint32_t n);
uint32_t hemk_spec_width(uint32_t L, int xf, uint32_t chunks);
/* Candidate-table entries (tab doubles, nx bytes) that launches of up to
 * max_rows chains of P chunks each can use, at the width each picks. */
size_t hemk_spec_tab_entries(uint32_t L, uint32_t max_rows, uint32_t P);
/* The allocation-free sums' fixed device scratch: chunk sums and drifts
 * (2 doubles), window base (8 bytes), candidates (HEMK_SPEC_W doubles) and
 * links (HEMK_SPEC_W bytes) per chunk, and two sums per row, for up to
 * HEMK_NOALLOC_CHUNKS chunks (rows x chunks per row), which also bounds the
 * rows.  hemk_noalloc_scratch: its address on the current device (looked
 * up on the first call on each device) and size. */
#define HEMK_NOALLOC_CHUNKS 4096
#define HEMK_NOALLOC_BYTES                                                   \
	((size_t)HEMK_NOALLOC_CHUNKS * (3 * 8 + HEMK_SPEC_W * 9 + 2 * 8))
#define HEMK_NOALLOC_DEVICES 64
int hemk_noalloc_scratch(void **ptr, size_t *bytes);
/* out[r] = sqrt(q[r] / (n - 1)) */
int hemk_sd_finish(const double *q, uint32_t rows, uint32_t n, double *out,
                   void *s);
/* produce == 1 (plain solve, nchain 2, rows 0/1 of x): first writes
 * r = u1 - u and v = (u2 - u1) - r (src/EM.c:333-336) into x. */
int hemk_spec_sum(const hemk_spec *sp, int produce, const double *u,
                  const double *u1, const double *u2, hemk_ctl *ctl,
                  void *s);

/* End of one iteration: k++, evaluate the loop condition of src/EM.c:328
 * from ctl->stopbad (the third map's stop test), clear it. */
int hemk_iter_end(hemk_ctl *ctl, void *s);

/* Finalisation (src/EM.c:362-373): est_t = lam_t len_t, tpm_t =
 * lam_t 1e6 / sum lam.  part: HEMK_NORM_BLOCKS doubles of scratch. */
int hemk_finalize(uint32_t n, const double *lam, const uint32_t *len,
                  double *part, double *tpm, double *est, void *s);
/* Same, with the sum taken from ctl->lam_sum (hemk_seq_norm phase 0). */
int hemk_finalize_seq(uint32_t n, const double *lam, const uint32_t *len,
                      const hemk_ctl *ctl, double *tpm, double *est, void *s);

/* FP64 log-likelihood; part: HEMK_NORM_BLOCKS*2 doubles; result in *out. */
int hemk_loglik(const hemk_prob *p, const uint32_t *unique, const double *lam,
                double *part, double *out, void *s);

#ifdef __cplusplus
}
#endif
#endif
