/*
 * solve.c -- the SQUAREM driver of one EM instance (src/EM.c:315-387).
 *
 * One iteration of the reference loop (src/EM.c:329-359) is captured once
 * into a HIP graph, HEM_ITERS_PER_GRAPH times over, and the graph is
 * replayed until the device-side control block reports that the loop
 * condition of src/EM.c:328 has failed.  Buffers rotate in a fixed pattern
 * so every iteration has the same pointers:
 *   A = u (iterate)   B = u1 = M(A)   C = u2 = M(B)
 *   D = extrapolated, clamped point   A <- M(D)  (stop test on this map)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <time.h>

#include "hem_internal.h"

void hem_opts_default(hem_opts *o)
{
	if (!o)
		return;
	memset(o, 0, sizeof(*o));
	o->precision = HEM_PREC_FP64;
	o->step = HEM_STEP_SD;
	o->clamp = HEM_CLAMP_ZERO;
	o->alpha_clamp = 1;
	o->norm = HEM_NORM_EXACT;
	o->min_iter = 50;
	o->max_iter = 1000;
	o->stop_min_value = 1e-3;
	o->stop_rel_change = 0.01;
	o->fallback = 0;
	o->verbose = 0;
}

static double now_ms(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static int check_opts(const hem_opts *o)
{
	if (o->precision != HEM_PREC_FP32 && o->precision != HEM_PREC_FP64)
		return hem_fail(HEM_ERR_INVALID, "unknown precision %d",
		                o->precision);
	if (o->step != HEM_STEP_SD && o->step != HEM_STEP_L2)
		return hem_fail(HEM_ERR_INVALID, "unknown step form %d", o->step);
	if (o->clamp != HEM_CLAMP_ZERO && o->clamp != HEM_CLAMP_DIV10)
		return hem_fail(HEM_ERR_INVALID, "unknown clamp form %d",
		                o->clamp);
	if (!hem_norm_exact(o->norm) && o->norm != HEM_NORM_TREE &&
	    o->norm != HEM_NORM_SEQ)
		return hem_fail(HEM_ERR_INVALID, "unknown norm mode %d", o->norm);
	if (o->fallback != 0 && o->fallback != 1)
		return hem_fail(HEM_ERR_INVALID, "fallback must be 0 or 1 (%d)",
		                o->fallback);
	return HEM_OK;
}

static int graph_key(const hem_problem *p, const hem_opts *o)
{
	return o->precision | (o->step << 1) | (o->clamp << 2) |
	       ((o->alpha_clamp != 0) << 3) | (o->norm << 4) |
	       (p->group_cls << 6) | (p->group_tx << 14) |
	       ((o->fallback != 0) << 24);
}

/* Optional per-kernel timing: an event is recorded after every launch. */
typedef struct {
	void *ev[16];
	int kind[16];
	int n;
} prof_ctx;

#define MARK(kd)                                                        \
	do {                                                            \
		if (pc) {                                               \
			HEM_TRY_DEV(hemd_event_record(pc->ev[pc->n], s),\
			            "event");                           \
			pc->kind[pc->n++] = (kd);                       \
		}                                                       \
	} while (0)

/* The exact-sum chains of the plain solve: rows of x (r/v rows, or the
 * state for the finalisation sum), results into res[row]. */
static hemk_spec spec_plain(const hem_problem *p, const hem_opts *o,
                            uint32_t nchain, int xf, const double *x,
                            const double *mres, double *res)
{
	hemk_spec sp;
	memset(&sp, 0, sizeof(sp));
	sp.n = p->n_tx;
	sp.L = p->spec_L;
	sp.P = p->spec_P;
	sp.nchain = nchain;
	sp.W = hemk_spec_width(sp.L, xf, nchain * sp.P);
	sp.R = 1;
	sp.xf = xf;
	sp.x = x;
	sp.ld = p->n_tx;
	sp.mres = mres;
	sp.csum = p->d_scsum;
	sp.drift = p->d_sdrift;
	sp.tab = p->d_stab;
	sp.base = p->d_sbase;
	sp.nx = p->d_snx;
	sp.res = res;
	sp.miss = &p->d_ctl->reruns;
	if (o->norm == HEM_NORM_EXACT) {
		sp.wa = p->d_swa;
		sp.wlh = p->d_swlh;
		sp.wrg = p->d_swrg;
		sp.wcnt = p->d_swcnt;
		sp.cpub = p->d_scpub;
	}
	return sp;
}

/* Solves run over the component tiles when every component fits a tile
 * (one kernel for u -> u1 -> u2, one for the extrapolated map): FP64, and
 * the FP32 bulk form on tables without member weights (the fallback is
 * FP64 only). */
static int use_tiles(const hem_problem *p, const hem_opts *o)
{
	size_t lds = hemk_ptile_lds(&p->dtp);
	const int f32 = o->precision == HEM_PREC_FP32;
	/* the fallback's kernel also keeps the class sums at u2 */
	if (o->fallback)
		lds = (lds + 7) / 8 * 8 + 8 * (size_t)p->dtp.max_cls;
	return p->dtp.n_tiles &&
	       (!f32 || (!p->weighted && !o->fallback)) &&
	       (hem_norm_exact(o->norm) || o->norm == HEM_NORM_TREE) &&
	       lds <= HEM_PTILE_LDS_MAX;
}

/* The first iteration's u1 = M(u), u2 = M(u1) over the tiles (and its r,
 * v rows for the exact sums); every later iteration's come from the
 * kernel that ends the iteration before it. */
static int enqueue_tiles_prologue(hem_problem *p, const hem_opts *o)
{
	const int exact = hem_norm_exact(o->norm);
	const uint32_t T = p->n_tx;
	HEM_TRY_DEV(hemk_ptile_map12(&p->dtp, p->d_A, p->d_B, p->d_C, p->d_uq,
	                             p->d_cntq, exact ? p->d_sx : NULL,
	                             exact ? p->d_sx + T : NULL,
	                             exact ? NULL : p->d_tpart, o->step,
	                             o->precision == HEM_PREC_FP32, p->d_ctl,
	                             p->stream),
	            "tile maps");
	if (!exact) {
		/* the first step length; each later one is folded by the
		 * kernel that ends the iteration before it */
		hemk_step st = {o->step, o->clamp, o->alpha_clamp != 0, 1};
		HEM_TRY_DEV(hemk_ptile_tree_alpha(&p->dtp, p->d_tpart, T, st, 0,
		                                  p->d_ctl, p->stream),
		            "alpha");
	}
	return HEM_OK;
}

/* One SQUAREM iteration over the tiles, its u1 and u2 already in B and C:
 * the norms, then one kernel that extrapolates, maps with the stop test
 * and maps twice more for the next iteration.  pos: the iteration's place
 * in a captured graph (-1: enqueued by itself).  With chunk programs the
 * iteration end of graph iterations but the last is done by the next
 * iteration's first walk (its kernels before that run regardless and are
 * never read once the loop has ended); the last one, and every iteration
 * enqueued by itself, ends with its own kernel. */
static int enqueue_iteration_tiles(hem_problem *p, const hem_opts *o,
                                   prof_ctx *pc, int pos)
{
	const int exact = hem_norm_exact(o->norm);
	const int sd = o->step == HEM_STEP_SD;
	const uint32_t T = p->n_tx;
	hemk_step st = {o->step, o->clamp, o->alpha_clamp != 0, exact ? 2 : 1};
	const double *alpha_part = NULL;
	void *s = p->stream;
	const int fuse_end = o->norm == HEM_NORM_EXACT && !o->fallback &&
	                     pos >= 0;

	if (o->fallback) {
		/* u1, u2 (and the r/v rows) at the start of every iteration,
		 * with the second map's stop test for an iteration that keeps
		 * u2 */
		HEM_TRY_DEV(hemk_ptile_map12_fb(&p->dtp, p->d_A, p->d_B, p->d_C,
		                                p->d_uq, p->d_cntq,
		                                exact ? p->d_sx : NULL,
		                                exact ? p->d_sx + T : NULL,
		                                exact ? NULL : p->d_tpart,
		                                o->step, p->d_ctl, s),
		            "tile maps");
		MARK(HEM_PK_TMAP12);
	}
	if (exact) {
		/* the reference's left-to-right sums over the r and v rows;
		 * with the sd form pass 1 also keeps each chunk's sum of
		 * squares, from which pass 2 (centred sums of squares) seeds
		 * its guesses without a chunk-sums step of its own */
		hemk_spec sp = spec_plain(p, o, 2,
		                          sd ? HEMK_SPEC_IDENT : HEMK_SPEC_SQUARE,
		                          p->d_sx, NULL, p->d_sres);
		if (sd)
			sp.csq = p->d_scsq;
		if (fuse_end && pos > 0)
			sp.iend = p->d_ctl;
		HEM_TRY_DEV(hemk_spec_sum(&sp, 0, NULL, NULL, NULL, p->d_ctl, s),
		            "exact sums 1");
		MARK(HEM_PK_NORM1);
		alpha_part = p->d_sres;
		if (sd) {
			sp = spec_plain(p, o, 2, HEMK_SPEC_CENTRED, p->d_sx,
			                p->d_sres, p->d_sres + 2);
			/* pass 1's chunk sums: its published words, or the chunk
			 * sums step's */
			sp.gsum = sp.cpub ? (const double *)p->d_scpub
			                  : p->d_scsum;
			sp.gsq = p->d_scsq;
			HEM_TRY_DEV(hemk_spec_phases(&sp,
			                             HEMK_SPEC_PH_DRIFT |
			                                 HEMK_SPEC_PH_FINISH,
			                             p->d_ctl, s),
			            "exact sums 2");
			MARK(HEM_PK_NORM2);
			alpha_part = p->d_sres + 2;
		}
	} else if (o->fallback) {
		/* tree norms: the tile kernel that made u1 and u2 left each
		 * tile's partials of r and v; fold them in tile order */
		HEM_TRY_DEV(hemk_ptile_tree_alpha(&p->dtp, p->d_tpart, T, st, 0,
		                                  p->d_ctl, s),
		            "alpha");
		MARK(HEM_PK_NORM1);
	}
	if (o->fallback) {
		/* extrapolate and stabilise; keep u2 instead when that lowers
		 * the log-likelihood */
		HEM_TRY_DEV(hemk_ptile_ex3_fb(&p->dtp, p->d_A, p->d_B, p->d_C,
		                              p->d_uq, p->d_cntq, alpha_part, T,
		                              st, p->d_fbpart, p->d_ctl, s),
		            "tile extrapolated map");
		MARK(HEM_PK_TEX3);
		HEM_TRY_DEV(hemk_fb_decide(p->d_fbpart, p->dtp.n_tiles,
		                           HEM_FB_TOL, p->d_ctl, s),
		            "fallback decision");
		HEM_TRY_DEV(hemk_fb_apply(T, p->d_C, p->d_A, p->d_ctl, s),
		            "fallback");
		MARK(HEM_PK_ITER_END);
	} else {
		HEM_TRY_DEV(hemk_ptile_ex3_map12(&p->dtp, p->d_A, p->d_B,
		                                 p->d_C, p->d_uq, p->d_cntq,
		                                 alpha_part, T, st,
		                                 exact ? p->d_sx : NULL,
		                                 exact ? p->d_sx + T : NULL,
		                                 exact ? NULL : p->d_tpart,
		                                 o->precision == HEM_PREC_FP32,
		                                 p->d_ctl, s),
		            "tile extrapolated map");
		MARK(HEM_PK_TEX3);
		if (!exact) {
			/* tree norms: end this iteration and fold the next
			 * step length from the partials just written */
			HEM_TRY_DEV(hemk_ptile_tree_alpha(&p->dtp, p->d_tpart, T,
			                                  st, 1, p->d_ctl, s),
			            "iteration end and alpha");
			MARK(HEM_PK_NORM1);
			return HEM_OK;
		}
	}
	if (!fuse_end || pos == HEM_ITERS_PER_GRAPH - 1) {
		HEM_TRY_DEV(hemk_iter_end(p->d_ctl, s), "iter_end");
		MARK(HEM_PK_ITER_END);
	}
	return HEM_OK;
}

/* Enqueue one SQUAREM iteration on the problem's stream (pos: its place in
 * a captured graph, -1 by itself). */
static int enqueue_iteration(hem_problem *p, const hem_opts *o, prof_ctx *pc,
                             int pos)
{
	if (use_tiles(p, o))
		return enqueue_iteration_tiles(p, o, pc, pos);
	{
	const int f32 = o->precision == HEM_PREC_FP32;
	const int prec = f32 ? 0 : 1;
	const void *gA = f32 ? (const void *)p->d_A32 : (const void *)p->d_A;
	const void *gB = f32 ? (const void *)p->d_B32 : (const void *)p->d_B;
	const void *gD = f32 ? (const void *)p->d_D32 : (const void *)p->d_D;
	void *rho = f32 ? p->d_rho : (void *)p->d_nd;
	const int exact = hem_norm_exact(o->norm);
	const int seq = o->norm == HEM_NORM_SEQ;
	hemk_step st = {o->step, o->clamp, o->alpha_clamp != 0,
	                seq ? 1 : exact ? 2 : 0};
	const double *alpha_part = o->step == HEM_STEP_SD ? p->d_part2
	                                                   : p->d_part1;
	const hemk_prob *dp = &p->dp;
	void *s = p->stream;
	int gc = p->group_cls, gt = p->group_tx;

	HEM_TRY_DEV(hemk_denom(dp, prec, gA, rho, gc, p->d_ctl, s),
	            "denom");
	MARK(HEM_PK_DENOM);
	HEM_TRY_DEV(hemk_gather_mstep(dp, prec, rho, p->d_unique, p->d_A,
	                              p->d_B, f32 ? p->d_B32 : NULL, NULL, gt,
	                              p->d_ctl, s),
	            "gather");
	MARK(HEM_PK_GATHER);
	HEM_TRY_DEV(hemk_denom(dp, prec, gB, rho, gc, p->d_ctl, s),
	            "denom");
	MARK(HEM_PK_DENOM);
	HEM_TRY_DEV(hemk_gather_mstep(dp, prec, rho, p->d_unique, p->d_B,
	                              p->d_C, NULL, NULL, gt, p->d_ctl, s),
	            "gather");
	MARK(HEM_PK_GATHER);
	if (exact) {
		/* the reference's left-to-right sums, by speculation: r and v
		 * rows with their sums (or sums of squares, L2 form), then
		 * the centred sums of the sd form; the extrapolation takes
		 * alpha from the last pair */
		const int sd = o->step == HEM_STEP_SD;
		hemk_spec sp = spec_plain(p, o, 2,
		                          sd ? HEMK_SPEC_IDENT : HEMK_SPEC_SQUARE,
		                          p->d_sx, NULL, p->d_sres);
		HEM_TRY_DEV(hemk_spec_sum(&sp, 1, p->d_A, p->d_B, p->d_C,
		                          p->d_ctl, s),
		            "exact sums 1");
		MARK(HEM_PK_NORM1);
		alpha_part = p->d_sres;
		if (sd) {
			sp = spec_plain(p, o, 2, HEMK_SPEC_CENTRED, p->d_sx,
			                p->d_sres, p->d_sres + 2);
			HEM_TRY_DEV(hemk_spec_sum(&sp, 0, NULL, NULL, NULL,
			                          p->d_ctl, s),
			            "exact sums 2");
			MARK(HEM_PK_NORM2);
			alpha_part = p->d_sres + 2;
		}
	} else if (seq) {
		/* the same sums on one lane each */
		if (o->step == HEM_STEP_SD) {
			HEM_TRY_DEV(hemk_seq_norm(p->n_tx, p->d_A, p->d_B,
			                          p->d_C, 1, st, p->d_ctl, s),
			            "seq norm 1");
			MARK(HEM_PK_NORM1);
			HEM_TRY_DEV(hemk_seq_norm(p->n_tx, p->d_A, p->d_B,
			                          p->d_C, 2, st, p->d_ctl, s),
			            "seq norm 2");
			MARK(HEM_PK_NORM2);
		} else {
			HEM_TRY_DEV(hemk_seq_norm(p->n_tx, p->d_A, p->d_B,
			                          p->d_C, 3, st, p->d_ctl, s),
			            "seq norm");
			MARK(HEM_PK_NORM1);
		}
	} else {
	HEM_TRY_DEV(hemk_norm_pass1(p->n_tx, p->d_A, p->d_B, p->d_C,
	                            p->d_part1, o->step, p->d_ctl, s),
	            "norm1");
	MARK(HEM_PK_NORM1);
	if (o->step == HEM_STEP_SD) {
		HEM_TRY_DEV(hemk_norm_pass2(p->n_tx, p->d_A, p->d_B, p->d_C,
		                            p->d_part1, p->d_part2, o->step,
		                            p->d_ctl, s),
		            "norm2");
		MARK(HEM_PK_NORM2);
	}
	}
	HEM_TRY_DEV(hemk_extrapolate(p->n_tx, p->d_A, p->d_B, p->d_C,
	                             alpha_part, p->d_D,
	                             f32 ? p->d_D32 : NULL, st, p->d_ctl, s),
	            "extrapolate");
	MARK(HEM_PK_EXTRAP);
	HEM_TRY_DEV(hemk_denom(dp, prec, gD, rho, gc, p->d_ctl, s),
	            "denom");
	MARK(HEM_PK_DENOM);
	HEM_TRY_DEV(hemk_gather_mstep(dp, prec, rho, p->d_unique, p->d_D,
	                              p->d_A, f32 ? p->d_A32 : NULL,
	                              &p->d_ctl->stopbad, gt, p->d_ctl, s),
	            "gather");
	MARK(HEM_PK_GATHER);
	HEM_TRY_DEV(hemk_iter_end(p->d_ctl, s), "iter_end");
	MARK(HEM_PK_ITER_END);
	return HEM_OK;
	}
}

/* Returns the iteration graph for these options, capturing it on first
 * use.  Graphs stay cached (no destroy while the handle lives). */
static int ensure_graph(hem_problem *p, const hem_opts *o, void **graph)
{
	int key = graph_key(p, o), rc, e;
	void *exec = NULL;
	for (int g = 0; g < p->n_graph; ++g)
		if (p->graph_key[g] == key) {
			*graph = p->graph[g];
			return HEM_OK;
		}
	if (p->n_graph == HEM_GRAPH_CACHE)
		return hem_fail(HEM_ERR_RANGE, "more than %d option sets on one "
		                "problem handle", HEM_GRAPH_CACHE);
	e = hemd_graph_begin(p->stream);
	if (e)
		return hem_dev_fail(e, "graph capture begin");
	for (int i = 0; i < HEM_ITERS_PER_GRAPH; ++i) {
		rc = enqueue_iteration(p, o, NULL, i);
		if (rc) {
			hemd_graph_end(p->stream, &exec);
			hemd_graph_destroy(exec);
			return rc;
		}
	}
	e = hemd_graph_end(p->stream, &exec);
	if (e)
		return hem_dev_fail(e, "graph capture end");
	p->graph[p->n_graph] = exec;
	p->graph_key[p->n_graph++] = key;
	*graph = exec;
	return HEM_OK;
}

/* Waits until the device has completed `target` iterations or ended the
 * loop, as it mirrors both into pinned memory; after 10 s without either the
 * stream is synchronised instead (a device error surfaces there). */
static int wait_iters(hem_problem *p, uint32_t target)
{
	volatile const uint32_t *h = p->h_done;
	struct timespec t0, t1;
	clock_gettime(CLOCK_MONOTONIC, &t0);
	for (uint32_t n = 1;; ++n) {
		if (h[0] || h[1] >= target)
			return HEM_OK;
		if ((n & 0xFFFu) == 0) {
			clock_gettime(CLOCK_MONOTONIC, &t1);
			if (t1.tv_sec - t0.tv_sec > 10) {
				const int e = hemd_stream_sync(p->stream);
				return e ? hem_dev_fail(e, "stream sync") : HEM_OK;
			}
		}
	}
}

/* Runs the iteration loop on the device from the state already in A.  From
 * min_iter on the host checks the done flag the device mirrors into pinned
 * memory one graph behind, with no copy or event on the stream: graph g + 1
 * is queued before the host waits for the device's iteration count to
 * reach the end of graph g, so the device never idles on the host's check;
 * the graph queued after convergence does nothing (every kernel reads the
 * done flag first). */
static int run_loop(hem_problem *p, const hem_opts *o, float *ms)
{
	int rc;
	void *graph = NULL;
	uint32_t launched = 0;
	int pending = 0;
	const uint32_t limit = o->max_iter > o->min_iter ? o->max_iter
	                                                  : o->min_iter;
	if ((rc = ensure_graph(p, o, &graph)))
		return rc;
	HEM_TRY_DEV(hemd_event_record(p->ev0, p->stream), "event");
	for (;;) {
		HEM_TRY_DEV(hemd_graph_launch(graph, p->stream),
		            "graph launch");
		launched += HEM_ITERS_PER_GRAPH;
		if (launched < o->min_iter && launched < limit)
			continue;
		if (pending) {
			/* the graph before this one has finished */
			if ((rc = wait_iters(p, launched - HEM_ITERS_PER_GRAPH)))
				return rc;
			if (((volatile uint32_t *)p->h_done)[0])
				break;
		}
		pending = 1;
		if (launched > limit + 2 * HEM_ITERS_PER_GRAPH)
			return hem_fail(HEM_ERR_DEVICE,
			                "iteration loop did not terminate");
	}
	HEM_TRY_DEV(hemd_d2h(p->h_ctl, p->d_ctl, sizeof(hemk_ctl), p->stream),
	            "control readback");
	HEM_TRY_DEV(hemd_event_record(p->ev1, p->stream), "event");
	HEM_TRY_DEV(hemd_event_sync(p->ev1), "event sync");
	HEM_TRY_DEV(hemd_event_elapsed(p->ev0, p->ev1, ms), "event time");
	return HEM_OK;
}

/* Uploads the unique counts, sets the uniform start (src/EM.c:311-312)
 * and resets the control block. */
static int solve_setup(hem_problem *p, const uint32_t *unique,
                       const hem_opts *o)
{
	int rc = hemd_h2d(p->d_unique, unique, (size_t)p->n_tx * 4, p->stream);
	/* tile order: unique counts by tile position, class counts by class
	 * tile position (the uniform start is the same in any order) */
	if (!rc && use_tiles(p, o))
		rc = hemk_gather_rows_u32(p->d_unique, p->dtp.tx, p->n_tx, 1,
		                          p->d_uq, p->stream);
	if (!rc && use_tiles(p, o))
		rc = hem_cntq_update(p);
	if (!rc)
		rc = hemk_init_state(p->n_tx, p->d_A, p->d_A32,
		                     (double)1 / p->n_tx, NULL, p->stream);
	((volatile uint32_t *)p->h_done)[0] = 0;
	((volatile uint32_t *)p->h_done)[1] = 0;
	if (!rc)
		rc = hemk_ctl_reset(p->d_ctl, o->min_iter, o->max_iter,
		                    o->stop_min_value, o->stop_rel_change,
		                    p->h_done, p->h_done + 1, p->stream);
	if (rc)
		return hem_dev_fail(rc, "solve setup");
	return use_tiles(p, o) && !o->fallback ? enqueue_tiles_prologue(p, o)
	                                       : HEM_OK;
}

static int load_opts(const hem_opts *opts, hem_opts *o)
{
	if (opts)
		*o = *opts;
	else
		hem_opts_default(o);
	return check_opts(o);
}

/* Member weights run on the FP64 order-faithful class and transcript
 * kernels only; the monotonicity fallback runs on the component tiles. */
static int check_weighted(const hem_problem *p, const hem_opts *o)
{
	if (o->fallback && !use_tiles(p, o))
		return hem_fail(HEM_ERR_RANGE,
		                "the monotonicity fallback runs on the component "
		                "tiles: FP64 with exact or tree norms, on a table "
		                "whose components fit a tile");
	if (p->weighted && o->precision != HEM_PREC_FP64)
		return hem_fail(HEM_ERR_INVALID, "a table with member weights "
		                                 "solves in FP64 only");
	return HEM_OK;
}

int hem_solve(hem_problem *p, const uint32_t *unique, const hem_opts *opts,
              double *tpm, double *est, double *lambda, hem_stats *st)
{
	hem_opts o;
	int rc;
	float ms = 0;
	double t0 = now_ms(), ll = 0;
	const double *lam;
	if (!p || !unique)
		return hem_fail(HEM_ERR_INVALID, "null problem or counts");
	if (p->ns > 1)
		return hem_fail(HEM_ERR_INVALID, "hem_solve needs a one-sample "
		                                 "handle (see hem_solve_batch)");
	if ((rc = load_opts(opts, &o)))
		return rc;
	if ((rc = check_weighted(p, &o)))
		return rc;

	pthread_mutex_lock(&p->mu);
	if ((rc = solve_setup(p, unique, &o)))
		goto out;
	if ((rc = run_loop(p, &o, &ms)))
		goto out;
	/* finalisation into B (est) and C (tpm); log-likelihood of the final
	 * state lam (A, or its transcript-order copy in D after a tiled
	 * solve) */
	lam = p->d_A;
	if (use_tiles(p, &o)) {
		rc = hemk_ptile_rows(p->n_tx, p->dtp.txq, p->d_A, p->d_D,
		                     p->stream);
		lam = p->d_D;
	}
	if (rc) {
	} else if (hem_norm_exact(o.norm)) {
		/* sum of the final state (src/EM.c:362-363) into lam_sum */
		hemk_spec sp = spec_plain(p, &o, 1, HEMK_SPEC_IDENT, lam, NULL,
		                          &p->d_ctl->lam_sum);
		rc = hemk_spec_sum(&sp, 0, NULL, NULL, NULL, NULL, p->stream);
		if (!rc)
			rc = hemk_finalize_seq(p->n_tx, lam, p->d_len, p->d_ctl,
			                       p->d_C, p->d_B, p->stream);
	} else if (o.norm == HEM_NORM_SEQ) {
		hemk_step st0 = {0, 0, 0, 0};
		rc = hemk_seq_norm(p->n_tx, p->d_A, NULL, NULL, 0, st0,
		                   p->d_ctl, p->stream);
		if (!rc)
			rc = hemk_finalize_seq(p->n_tx, p->d_A, p->d_len, p->d_ctl,
			                       p->d_C, p->d_B, p->stream);
	} else {
		rc = hemk_finalize(p->n_tx, lam, p->d_len, p->d_part1,
		                   p->d_C, p->d_B, p->stream);
	}
	if (!rc && !p->weighted)
		rc = hemk_loglik(&p->dp, p->d_unique, lam, p->d_part2,
		                 p->d_out, p->stream);
	if (!rc && tpm)
		rc = hemd_d2h(tpm, p->d_C, (size_t)p->n_tx * 8, p->stream);
	if (!rc && est)
		rc = hemd_d2h(est, p->d_B, (size_t)p->n_tx * 8, p->stream);
	if (!rc && lambda)
		rc = hemd_d2h(lambda, lam, (size_t)p->n_tx * 8, p->stream);
	if (!rc)
		rc = hemd_d2h(p->h_ctl, p->d_ctl, sizeof(hemk_ctl), p->stream);
	if (!rc && !p->weighted)
		rc = hemd_d2h(&ll, p->d_out, sizeof(double), p->stream);
	if (p->weighted)
		ll = NAN;
	if (!rc)
		rc = hemd_stream_sync(p->stream);
	if (rc) {
		rc = hem_dev_fail(rc, "finalisation");
		goto out;
	}
	if (st) {
		memset(st, 0, sizeof(*st));
		st->iters = p->h_ctl->k;
		st->maps = p->h_ctl->maps;
		st->converged = p->h_ctl->converged;
		st->fallbacks = p->h_ctl->fallbacks;
		st->sum_reruns = p->h_ctl->reruns;
		st->last_alpha = p->h_ctl->alpha;
		st->loglik = ll;
		st->ms_solve = ms;
		st->ms_total = now_ms() - t0;
	}
	if (o.verbose)
		fprintf(stderr, "Finish EM with %u rounds\n", p->h_ctl->k);
out:
	pthread_mutex_unlock(&p->mu);
	return rc;
}

/* Algorithmic bytes per launch of each kernel kind: compulsory traffic with
 * every array element touched once (gathered vectors counted once). */
static void kernel_bytes(const hem_problem *p, int f32, int rv, double *b)
{
	const double T = p->n_tx, C = p->n_cls, Z = (double)p->nnz;
	if (f32) {
		/* counts, densities, rho = n/D all FP32 */
		b[HEM_PK_DENOM] = 4 * (C + 1) + 4 * Z + 4 * C + 4 * T + 4 * C;
		b[HEM_PK_GATHER] = 4 * (T + 1) + 4 * Z + 4 * C /* rho */ +
		                   4 * T /* unique */ + 4 * T /* len */ +
		                   8 * T /* lam */ + 8 * T /* out */ +
		                   4 * T /* FP32 shadow */;
	} else {
		/* densities FP64, D_c written, {n_c, D_c} pairs gathered */
		b[HEM_PK_DENOM] = 4 * (C + 1) + 4 * Z + 8 * T + 8 * C;
		b[HEM_PK_GATHER] = 4 * (T + 1) + 4 * Z + 16 * C +
		                   4 * T + 4 * T + 8 * T + 8 * T;
	}
	if (p->weighted) {
		/* one FP64 weight per member in each pass */
		b[HEM_PK_DENOM] += 8 * Z;
		b[HEM_PK_GATHER] += 8 * Z;
	}
	b[HEM_PK_NORM1] = 24 * T;   /* u, u1, u2 */
	b[HEM_PK_NORM2] = 24 * T;
	b[HEM_PK_EXTRAP] = 24 * T + 8 * T + (f32 ? 4 * T : 0);
	b[HEM_PK_ITER_END] = 8;
	{
		/* tile kernels: each tile's descriptor and record (offsets,
		 * members, occurrence positions, lengths, ids) once, its
		 * unique and class counts, then the state */
		const double NT = p->dtp.n_tiles, TC = p->dtp.n_cls;
		const double S = 32 * NT + (double)p->ptile_rec_bytes + 4 * T +
		                 8 * TC;
		b[HEM_PK_TMAP12] = S + 8 * T + 16 * T + 16 * T;
		/* A, B, C in; lam, u1, u2 out; r, v rows with the exact sums,
		 * or each tile's norm partials (the tree norms fold those:
		 * 32 bytes and a transcript count per tile) */
		b[HEM_PK_TEX3] = S + 24 * T + 24 * T + (rv ? 16 * T : 32 * NT);
		if (!rv && p->dtp.n_tiles)
			b[HEM_PK_NORM1] = 36 * NT;
	}
}

int hem_profile_solve(hem_problem *p, const uint32_t *unique,
                      const hem_opts *opts, hem_kernel_profile *out)
{
	hem_opts o;
	prof_ctx pc;
	void *ev_start = NULL;
	int rc, i;
	if (!p || !unique || !out)
		return hem_fail(HEM_ERR_INVALID, "null argument");
	if (p->ns > 1)
		return hem_fail(HEM_ERR_INVALID, "one-sample handles only");
	if ((rc = load_opts(opts, &o)))
		return rc;
	if ((rc = check_weighted(p, &o)))
		return rc;
	memset(out, 0, sizeof(*out));
	memset(&pc, 0, sizeof(pc));
	kernel_bytes(p, o.precision == HEM_PREC_FP32,
	             hem_norm_exact(o.norm), out->bytes);
	for (i = 0; i < 16 && !rc; ++i)
		rc = hemd_event_create(&pc.ev[i]);
	if (!rc)
		rc = hemd_event_create(&ev_start);
	if (rc) {
		rc = hem_dev_fail(rc, "event creation");
		goto free_ev;
	}
	pthread_mutex_lock(&p->mu);
	if ((rc = solve_setup(p, unique, &o)))
		goto out;
	for (;;) {
		int e;
		pc.n = 0;
		e = hemd_event_record(ev_start, p->stream);
		if (e) {
			rc = hem_dev_fail(e, "event");
			goto out;
		}
		if ((rc = enqueue_iteration(p, &o, &pc, -1)))
			goto out;
		e = hemd_d2h(p->h_ctl, p->d_ctl, sizeof(hemk_ctl), p->stream);
		if (!e)
			e = hemd_stream_sync(p->stream);
		if (e) {
			rc = hem_dev_fail(e, "profiled iteration");
			goto out;
		}
		for (i = 0; i < pc.n; ++i) {
			float ms = 0;
			e = hemd_event_elapsed(i ? pc.ev[i - 1] : ev_start, pc.ev[i],
			                       &ms);
			if (e) {
				rc = hem_dev_fail(e, "event time");
				goto out;
			}
			out->ms[pc.kind[i]] += ms;
			out->launches[pc.kind[i]]++;
		}
		out->iters++;
		if (p->h_ctl->done)
			break;
		if (out->iters > o.max_iter + o.min_iter + 1) {
			rc = hem_fail(HEM_ERR_DEVICE,
			              "profiled loop did not terminate");
			goto out;
		}
	}
	out->iters = p->h_ctl->k;
out:
	pthread_mutex_unlock(&p->mu);
free_ev:
	for (i = 0; i < 16; ++i)
		if (pc.ev[i])
			hemd_event_destroy(pc.ev[i]);
	if (ev_start)
		hemd_event_destroy(ev_start);
	return rc;
}

int hem_loglik(hem_problem *p, const uint32_t *unique, const double *lambda,
               double *loglik)
{
	int rc;
	if (!p || !unique || !lambda || !loglik)
		return hem_fail(HEM_ERR_INVALID, "null argument");
	if (p->ns > 1)
		return hem_fail(HEM_ERR_INVALID, "one-sample handles only");
	if (p->weighted)
		return hem_fail(HEM_ERR_INVALID, "no log-likelihood for a "
		                                 "table with member weights");
	pthread_mutex_lock(&p->mu);
	rc = hemd_h2d(p->d_unique, unique, (size_t)p->n_tx * 4, p->stream);
	if (!rc)
		rc = hemd_h2d(p->d_D, lambda, (size_t)p->n_tx * 8, p->stream);
	if (!rc)
		rc = hemk_loglik(&p->dp, p->d_unique, p->d_D, p->d_part2,
		                 p->d_out, p->stream);
	if (!rc)
		rc = hemd_d2h(loglik, p->d_out, sizeof(double), p->stream);
	if (!rc)
		rc = hemd_stream_sync(p->stream);
	pthread_mutex_unlock(&p->mu);
	return rc ? hem_dev_fail(rc, "log-likelihood") : HEM_OK;
}

/* Exact left-to-right sums of host rows on the device (the reductions of
 * Norm, src/hash_align.c:150-164, and src/EM.c:362-363). */
int hem_exact_sums(const double *x, uint32_t n, uint32_t n_rows, int mode,
                   double *out, uint32_t *reruns)
{
	int rc, e = 0;
	void *s = NULL;
	double *d_x = NULL, *d_sc = NULL, *d_tab = NULL, *d_res = NULL;
	uint8_t *d_nx = NULL;
	long long *d_base = NULL;
	uint32_t *d_miss = NULL, miss = 0;
	double *d_wa = NULL, *d_wlh = NULL;
	uint32_t *d_wrg = NULL, *d_wcnt = NULL;
	unsigned long long *d_cpub = NULL;
	hemk_spec sp;
	const int windows = (mode & HEM_SUM_WINDOWS) != 0;
	mode &= ~HEM_SUM_WINDOWS;
	if (!x || !out || n == 0 || n_rows == 0)
		return hem_fail(HEM_ERR_INVALID, "null rows or empty size");
	if (mode < HEM_SUM_PLAIN || mode > HEM_SUM_SQUARES)
		return hem_fail(HEM_ERR_INVALID, "unknown sum mode %d", mode);
	if (n_rows > 65535 || n > 0xFFFFFFFFu - HEMK_SPEC_LMAX)
		return hem_fail(HEM_ERR_RANGE, "at most 65535 rows of < 2^32 "
		                "values");
	if ((rc = hem_device_ready()))
		return rc;
	memset(&sp, 0, sizeof(sp));
	sp.n = n;
	sp.L = hemk_spec_chunk(n);
	sp.P = (n + sp.L - 1) / sp.L;
	sp.nchain = n_rows;
	sp.R = 1;
	sp.ld = n;
	const size_t cp = (size_t)n_rows * sp.P;
	if (!e) e = hemd_stream_create(&s);
	if (!e) e = hemd_malloc((void **)&d_x, (size_t)n_rows * n * 8);
	if (!e) e = hemd_malloc((void **)&d_sc, cp * 2 * 8);
	if (!e) e = hemd_malloc((void **)&d_base, cp * 8);
	if (!e) e = hemd_malloc((void **)&d_tab, cp * HEMK_SPEC_W * 8);
	if (!e) e = hemd_malloc((void **)&d_nx, cp * HEMK_SPEC_W);
	if (!e) e = hemd_malloc((void **)&d_res, (size_t)n_rows * 2 * 8);
	if (!e) e = hemd_malloc((void **)&d_miss, 4);
	if (!windows) {
		const size_t nrec = hemk_ws_records(sp.L, n_rows, sp.P);
		if (!e) e = hemd_malloc((void **)&d_wa, nrec * 8);
		if (!e) e = hemd_malloc((void **)&d_wlh, nrec * 16);
		if (!e) e = hemd_malloc((void **)&d_wrg, nrec * 4);
		if (!e) e = hemd_malloc((void **)&d_wcnt, cp * 4);
		if (!e) e = hemd_malloc((void **)&d_cpub, cp * 8);
	}
	if (e) {
		rc = hem_fail(HEM_ERR_NOMEM, "device allocation: %s",
		              hemd_error_string(e));
		goto out;
	}
	sp.x = d_x;
	sp.csum = d_sc;
	sp.drift = d_sc + cp;
	sp.tab = d_tab;
	sp.base = d_base;
	sp.nx = d_nx;
	sp.miss = d_miss;
	sp.res = d_res;
	sp.wa = d_wa;
	sp.wlh = d_wlh;
	sp.wrg = d_wrg;
	sp.wcnt = d_wcnt;
	sp.cpub = d_cpub;
	sp.xf = mode == HEM_SUM_SQUARES ? HEMK_SPEC_SQUARE : HEMK_SPEC_IDENT;
	sp.W = hemk_spec_width(sp.L, sp.xf, sp.nchain * sp.P);
	e = hemd_h2d(d_x, x, (size_t)n_rows * n * 8, s);
	if (!e) e = hemk_fill_u32(d_miss, 1, 0, s);
	if (!e && d_cpub)
		e = hemk_fill_u32((uint32_t *)d_cpub, 2 * cp, 0, s);
	if (!e) e = hemk_spec_sum(&sp, 0, NULL, NULL, NULL, NULL, s);
	if (!e && mode == HEM_SUM_NORM_SD) {
		/* second pass of Norm: centred squares about sum / n */
		sp.xf = HEMK_SPEC_CENTRED;
		sp.W = hemk_spec_width(sp.L, sp.xf, sp.nchain * sp.P);
		sp.mres = d_res;
		sp.res = d_res + n_rows;
		e = hemk_spec_sum(&sp, 0, NULL, NULL, NULL, NULL, s);
	}
	if (!e) e = hemd_d2h(out, mode == HEM_SUM_NORM_SD ? d_res + n_rows
	                                                 : d_res,
	                     (size_t)n_rows * 8, s);
	if (!e) e = hemd_d2h(&miss, d_miss, 4, s);
	if (!e) e = hemd_stream_sync(s);
	if (e) {
		rc = hem_dev_fail(e, "exact sums");
		goto out;
	}
	if (mode == HEM_SUM_NORM_SD)
		for (uint32_t r = 0; r < n_rows; ++r)
			out[r] = sqrt(out[r] / (n - 1));
	if (reruns)
		*reruns = miss;
	rc = HEM_OK;
out:
	hemd_free(d_x);
	hemd_free(d_sc);
	hemd_free(d_base);
	hemd_free(d_tab);
	hemd_free(d_nx);
	hemd_free(d_res);
	hemd_free(d_miss);
	hemd_free(d_wa);
	hemd_free(d_wlh);
	hemd_free(d_wrg);
	hemd_free(d_wcnt);
	hemd_free(d_cpub);
	if (s)
		hemd_stream_destroy(s);
	return rc;
}

/* The sums of hem_exact_sums over device rows, on the caller's stream, with
 * the library's fixed device scratch instead of allocations. */
int hem_exact_sums_noalloc(const double *x, uint32_t n, uint32_t n_rows,
                           int mode, double *out, uint32_t *reruns,
                           void *stream)
{
	hemk_spec sp;
	void *mem = NULL;
	size_t bytes = 0, cp;
	uint32_t L, P;
	double *csum, *drift, *res, *tab;
	long long *base;
	uint8_t *nx;
	int rc, e;
	if (!x || !out || n == 0 || n_rows == 0)
		return hem_fail(HEM_ERR_INVALID, "null rows or empty size");
	if (mode < HEM_SUM_PLAIN || mode > HEM_SUM_SQUARES)
		return hem_fail(HEM_ERR_INVALID, "unknown sum mode %d", mode);
	if (n > 0xFFFFFFFFu - HEMK_SPEC_LMAX)
		return hem_fail(HEM_ERR_RANGE, "rows of < 2^32 values");
	L = hemk_spec_chunk(n);
	P = (n + L - 1) / L;
	cp = (size_t)n_rows * P;
	if (cp > HEMK_NOALLOC_CHUNKS)
		return hem_fail(HEM_ERR_RANGE,
		                "%u rows of %u values are %zu chunks of %u; the "
		                "allocation-free sums take at most %d chunks",
		                n_rows, n, cp, L, HEMK_NOALLOC_CHUNKS);
	if ((rc = hem_device_ready()))
		return rc;
	e = hemk_noalloc_scratch(&mem, &bytes);
	if (e)
		return hem_dev_fail(e, "allocation-free scratch");
	if (bytes < HEMK_NOALLOC_BYTES)
		return hem_fail(HEM_ERR_DEVICE, "scratch of %zu bytes", bytes);
	/* [C] chunk sums, [C] drifts, [C] bases, [2C] sums by row, [C][W]
	 * candidates, [C][W] links; C = HEMK_NOALLOC_CHUNKS */
	csum = (double *)mem;
	drift = csum + HEMK_NOALLOC_CHUNKS;
	base = (long long *)(drift + HEMK_NOALLOC_CHUNKS);
	res = (double *)(base + HEMK_NOALLOC_CHUNKS);
	tab = res + 2 * (size_t)HEMK_NOALLOC_CHUNKS;
	nx = (uint8_t *)(tab + (size_t)HEMK_NOALLOC_CHUNKS * HEMK_SPEC_W);
	memset(&sp, 0, sizeof(sp));
	sp.n = n;
	sp.L = L;
	sp.P = P;
	sp.nchain = n_rows;
	sp.R = 1;
	sp.ld = n;
	sp.x = x;
	sp.csum = csum;
	sp.drift = drift;
	sp.tab = tab;
	sp.base = base;
	sp.nx = nx;
	sp.miss = reruns;
	sp.xf = mode == HEM_SUM_SQUARES ? HEMK_SPEC_SQUARE : HEMK_SPEC_IDENT;
	sp.W = hemk_spec_width(L, sp.xf, (uint32_t)cp);
	sp.res = mode == HEM_SUM_NORM_SD ? res : out;
	e = reruns ? hemk_fill_u32(reruns, 1, 0, stream) : 0;
	if (!e)
		e = hemk_spec_sum(&sp, 0, NULL, NULL, NULL, NULL, stream);
	if (!e && mode == HEM_SUM_NORM_SD) {
		/* second pass of Norm: centred squares about sum / n, then
		 * sqrt(q / (n - 1)) */
		sp.xf = HEMK_SPEC_CENTRED;
		sp.W = hemk_spec_width(L, sp.xf, (uint32_t)cp);
		sp.mres = res;
		sp.res = res + n_rows;
		e = hemk_spec_sum(&sp, 0, NULL, NULL, NULL, NULL, stream);
		if (!e)
			e = hemk_sd_finish(res + n_rows, n_rows, n, out, stream);
	}
	return e ? hem_dev_fail(e, "allocation-free exact sums") : HEM_OK;
}
