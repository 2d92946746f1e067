/*
 * batch.c -- many EM instances over one class table, and hera's bootstrap.
 *
 * The reference runs each bootstrap replicate as its own estimate_count in a
 * pthread (src/EM.c:123-150, batches of nThread joined at src/EM.c:243-255).
 * Here all replicates run together: the state is [transcript][replicate]
 * (replicate fastest), one workgroup per (component tile, replicate chunk)
 * reads the tile's structure once for its whole chunk, and one captured graph
 * iteration advances every replicate that is still iterating.  Each replicate
 * keeps hera's own loop: its k, its step length, its stop test
 * (src/EM.c:325-360).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "hem_internal.h"

static double now_ms(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* ---- tiles on the device ------------------------------------------------ */

int hem_tiles_upload(hem_problem *p, const hem_tiles_host *h, hemk_tiles *dt,
                     uint32_t **mem, uint32_t **cls_orig,
                     uint64_t *rec_bytes)
{
	void *s = p->stream;
	const size_t nz = h->nnz ? h->nnz : 1;
	size_t sz[HEM_TILE_ARRAYS] = {
	    (size_t)h->n_tiles * HEM_TILE_DESC * 4, (size_t)h->n_tx * 4,
	    (size_t)h->n_tx * 4, ((size_t)h->n_cls + 1) * 4, nz * 4,
	    ((size_t)h->n_tx + 1) * 4, nz * 2,
	    (size_t)(h->n_cls ? h->n_cls : 1) * 4, (size_t)h->n_tx * 4,
	    (size_t)(h->rec_units ? h->rec_units : 1) * 16};
	const void *src[HEM_TILE_ARRAYS] = {
	    h->desc, h->tx,   h->len,      h->coff, h->kmem,
	    h->toff, h->tcls, h->cls_orig, h->txq,  h->rec};
	for (int i = 0; i < HEM_TILE_ARRAYS; ++i) {
		int e = hemd_malloc((void **)&mem[i], sz[i]);
		if (e)
			return hem_fail(HEM_ERR_NOMEM, "tile upload: %s",
			                hemd_error_string(e));
		p->dev_bytes += sz[i];
		e = hemd_h2d(mem[i], src[i], sz[i], s);
		if (e)
			return hem_dev_fail(e, "tile upload");
	}
	dt->n_tiles = h->n_tiles;
	dt->n_tx = h->n_tx;
	dt->n_cls = h->n_cls;
	dt->max_tx = h->max_tx;
	dt->max_cls = h->max_cls;
	dt->max_nnz = h->max_nnz;
	dt->desc = mem[0];
	dt->tx = mem[1];
	dt->len = mem[2];
	dt->coff = mem[3];
	dt->kmem = mem[4];
	dt->toff = mem[5];
	dt->tcls = (const uint16_t *)mem[6];
	*cls_orig = mem[7];
	dt->txq = mem[8];
	dt->rec = mem[9];
	dt->max_rec = h->max_rec;
	dt->weighted = h->wcls != NULL;
	if (rec_bytes)
		*rec_bytes = h->rec_units * 16;
	return HEM_OK;
}

int hem_batch_chunk(const hem_problem *p, uint32_t R)
{
	int m = hemk_tile_chunk(&p->dt, R, 4);
	if (p->cfg.chunk && (int)p->cfg.chunk < m)
		m = (int)p->cfg.chunk;
	return m;
}

int hem_problem_tiles(const hem_problem *p, uint32_t *n_tiles,
                      uint32_t *max_tx, uint32_t *max_cls)
{
	if (!p)
		return hem_fail(HEM_ERR_INVALID, "null problem");
	if (n_tiles)
		*n_tiles = p->dt.n_tiles;
	if (max_tx)
		*max_tx = p->dt.max_tx;
	if (max_cls)
		*max_cls = p->dt.max_cls;
	return HEM_OK;
}

/* ---- workspace ---------------------------------------------------------- */

static void ws_free_graphs(hem_batch_ws *w)
{
	for (int g = 0; g < w->n_graph; ++g)
		hemd_graph_destroy(w->graph[g]);
	w->n_graph = 0;
}

/* Takes the cached batch graphs out of use without destroying them (they
 * hold the old buffers' addresses; destruction waits for the handle's). */
static int ws_retire_graphs(hem_problem *p)
{
	hem_batch_ws *w = &p->bw;
	if (p->n_retired + w->n_graph > p->cap_retired) {
		int cap = 2 * (p->n_retired + w->n_graph) + 8;
		void **nr = realloc(p->retired, (size_t)cap * sizeof(void *));
		if (!nr)
			return hem_fail(HEM_ERR_NOMEM, "host allocation");
		p->retired = nr;
		p->cap_retired = cap;
	}
	for (int g = 0; g < w->n_graph; ++g)
		p->retired[p->n_retired++] = w->graph[g];
	w->n_graph = 0;
	return HEM_OK;
}

void hem_batch_destroy_retired(hem_problem *p)
{
	for (int g = 0; g < p->n_retired; ++g)
		hemd_graph_destroy(p->retired[g]);
	free(p->retired);
	p->retired = NULL;
	p->n_retired = p->cap_retired = 0;
}

void hem_batch_free(hem_problem *p)
{
	hem_batch_ws *w = &p->bw;
	ws_free_graphs(w);
	hemd_free(w->A);
	hemd_free(w->B);
	hemd_free(w->C);
	hemd_free(w->cat_rm);
	hemd_free(w->catT);
	hemd_free(w->uq);
	hemd_free(w->cnt);
	hemd_free(w->part);
	hemd_free(w->X);
	hemd_free(w->m);
	hemd_free(w->rs_plan);
	hemd_free(w->seeds);
	hemd_free(w->u32);
	hemd_free(w->f64);
	hemd_free(w->sc);
	hemd_free(w->sbase);
	hemd_free(w->stab);
	hemd_free(w->snx);
	hemd_free(w->swa);
	hemd_free(w->swlh);
	hemd_free(w->swrg);
	hemd_free(w->swcnt);
	hemd_free(w->scpub);
	hemd_free(w->scsq);
	hemd_host_free(w->h_active);
	memset(w, 0, sizeof(*w));
	for (int i = 0; i < HEM_TILE_ARRAYS; ++i) {
		hemd_free(p->d_tile_mem[i]);
		p->d_tile_mem[i] = NULL;
	}
}

static int ws_ensure(hem_problem *p, uint32_t R)
{
	hem_batch_ws *w = &p->bw;
	const size_t T = p->n_tx, K = (size_t)p->n_tx + p->n_cls;
	const size_t NC = p->dt.n_cls ? p->dt.n_cls : 1;
	const size_t NT = p->dt.n_tiles ? p->dt.n_tiles : 1;
	if (w->R >= R)
		return HEM_OK;
	if (ws_retire_graphs(p))
		return HEM_ERR_NOMEM;
	{
		/* keep the tile arrays; drop the per-R buffers */
		void *keep[HEM_TILE_ARRAYS];
		memcpy(keep, p->d_tile_mem, sizeof(keep));
		memset(p->d_tile_mem, 0, sizeof(keep));
		hem_batch_free(p);
		memcpy(p->d_tile_mem, keep, sizeof(keep));
	}
#define WALLOC(ptr, bytes)                                                  \
	do {                                                                \
		int e_ = hemd_malloc((void **)&(ptr), (bytes));             \
		if (e_) {                                                   \
			return hem_fail(HEM_ERR_NOMEM,                      \
			                "batch workspace (%zu bytes): %s",  \
			                (size_t)(bytes),                    \
			                hemd_error_string(e_));             \
		}                                                           \
	} while (0)
	WALLOC(w->A, T * R * 8);
	WALLOC(w->B, T * R * 8);
	WALLOC(w->C, T * R * 8);
	WALLOC(w->cat_rm, K * R * 4);
	WALLOC(w->catT, K * R * 4);
	WALLOC(w->uq, T * R * 4);
	WALLOC(w->cnt, NC * R * 4);
	WALLOC(w->part, NT * R * 4 * 8);
	WALLOC(w->X, 2 * T * R * 8);
	WALLOC(w->m, (size_t)R * HEMK_RS_MAXB * 4);
	WALLOC(w->rs_plan, (K + 2 * ((size_t)HEMK_RS_MAXB + 1)) * 4);
	WALLOC(w->seeds, (size_t)R * 8);
	{
		const size_t ns = p->ns, NI = ns * R;
		WALLOC(w->u32, (4 * NI + ns + ns * ((size_t)R + 1) + 1) * 4);
		WALLOC(w->f64, 8 * NI * 8);
	}
	WALLOC(w->sc, (size_t)4 * R * p->spec_P * 8);
	WALLOC(w->sbase, (size_t)2 * R * p->spec_P * 8);
	{
		/* r/v pairs (2R rows) and the final sums (R rows), each at
		 * the window width its launch picks */
		const size_t ent = hemk_spec_tab_entries(p->spec_L, 2 * R,
		                                         p->spec_P);
		WALLOC(w->stab, ent * 8);
		WALLOC(w->snx, ent);
	}
	{
		/* chunk programs of the r/v pairs (the final sums' R rows use
		 * the first R chains) */
		const size_t nrec = hemk_ws_records(p->spec_L, 2 * R, p->spec_P);
		WALLOC(w->swa, nrec * 8);
		WALLOC(w->swlh, nrec * 16);
		WALLOC(w->swrg, nrec * 4);
		WALLOC(w->swcnt, (size_t)2 * R * p->spec_P * 4);
		WALLOC(w->scpub, (size_t)2 * R * p->spec_P * 8);
		WALLOC(w->scsq, (size_t)2 * R * p->spec_P * 8);
		if (hemk_fill_u32((uint32_t *)w->scpub,
		                  (size_t)4 * R * p->spec_P, 0, p->stream))
			return hem_fail(HEM_ERR_DEVICE, "batch workspace fill");
	}
#undef WALLOC
	if (hemd_host_alloc((void **)&w->h_active,
	                    (size_t)(p->ns > 16 ? p->ns : 16) * 4))
		return hem_fail(HEM_ERR_NOMEM, "pinned allocation");
	w->R = R;
	return HEM_OK;
}

/* Control arrays of the workspace, for NIc = ns * capacity instances:
 *   u32: k, done, conv, notconv [NIc each], active [ns],
 *        alist [ns][capacity + 1], exact-sum reruns [1]
 *   f64: alpha, mean_r, mean_v, lam_sum [NIc each], q_r/q_v [2 NIc],
 *        exact pass-1 sums [2 NIc] */
static uint32_t *reruns_word(const hem_problem *p)
{
	const size_t ns = p->ns, NIc = ns * p->bw.R;
	return p->bw.u32 + 4 * NIc + ns + ns * ((size_t)p->bw.R + 1);
}

static double *pass1_sums(const hem_problem *p)
{
	return p->bw.f64 + 6 * (size_t)p->ns * p->bw.R;
}

static hemk_rctl rctl_for(hem_problem *p, uint32_t R, const hem_opts *o)
{
	hem_batch_ws *w = &p->bw;
	const size_t NIc = (size_t)p->ns * w->R;
	hemk_rctl rc;
	memset(&rc, 0, sizeof(rc));
	rc.R = R;
	rc.min_iter = o->min_iter;
	rc.max_iter = o->max_iter;
	rc.stop_min = o->stop_min_value;
	rc.stop_rel = o->stop_rel_change;
	rc.k = w->u32;
	rc.done = w->u32 + NIc;
	rc.conv = w->u32 + 2 * NIc;
	rc.notconv = w->u32 + 3 * NIc;
	rc.active = w->u32 + 4 * NIc;
	rc.alist = w->u32 + 4 * NIc + p->ns;
	rc.alpha = w->f64;
	rc.mean_r = w->f64 + NIc;
	rc.mean_v = w->f64 + 2 * NIc;
	rc.lam_sum = w->f64 + 3 * NIc;
	/* q_v follows q_r at distance ns * R: the exact sums write both as
	 * one array of 2R rows (one sample) */
	rc.q_r = w->f64 + 4 * NIc;
	rc.q_v = rc.q_r + (size_t)p->ns * R;
	rc.ns = p->ns;
	if (p->ns > 1) {
		rc.tsamp = p->d_tsamp;
		rc.txsamp = p->d_txsamp;
		rc.sinfo = p->d_sinfo;
	}
	return rc;
}

/* Exact-sum chains over the batch's rows.  pairs: r and v rows of the
 * replicates in alist (X[rep] and X[R + rep]); otherwise one row per
 * replicate (every replicate, no alist). */
static hemk_spec spec_batch(hem_problem *p, const hem_opts *o,
                            hemk_rctl rc, int pairs, int xf,
                            const double *mres, double *res)
{
	hem_batch_ws *w = &p->bw;
	const size_t cap = w->R;
	hemk_spec sp;
	memset(&sp, 0, sizeof(sp));
	sp.n = p->n_tx;
	sp.L = p->spec_L;
	sp.P = p->spec_P;
	sp.nchain = pairs ? 2 * rc.R : rc.R;
	sp.W = hemk_spec_width(sp.L, xf, sp.nchain * sp.P);
	sp.R = rc.R;
	sp.pairs = pairs;
	sp.xf = xf;
	sp.x = w->X;
	sp.ld = p->n_tx;
	sp.alist = pairs ? rc.alist : NULL;
	sp.mres = mres;
	sp.csum = w->sc;
	sp.drift = w->sc + 2 * cap * p->spec_P;
	sp.tab = w->stab;
	sp.base = w->sbase;
	sp.nx = w->snx;
	sp.res = res;
	sp.miss = reruns_word(p);
	if (o->norm == HEM_NORM_EXACT) {
		sp.wa = w->swa;
		sp.wlh = w->swlh;
		sp.wrg = w->swrg;
		sp.wcnt = w->swcnt;
		sp.cpub = w->scpub;
	}
	return sp;
}

/* ---- the iteration ------------------------------------------------------ */

/* Optional per-launch timing of the batched loop: an event after every
 * launch (hem_profile_batch). */
typedef struct {
	void *ev[16];
	int kind[16];
	int n;
} bprof;

#define BMARK(kd)                                                       \
	do {                                                            \
		if (pc) {                                               \
			HEM_TRY_DEV(hemd_event_record(pc->ev[pc->n],    \
			                              p->stream),       \
			            "event");                           \
			pc->kind[pc->n++] = (kd);                       \
		}                                                       \
	} while (0)

/* The first iteration's u1, u2 and r/v rows (or tree partials); every
 * later iteration's come from the kernel that ends the one before it. */
static int batch_prologue(hem_problem *p, const hem_opts *o, hemk_rctl rc,
                          int chunk, bprof *pc)
{
	hem_batch_ws *w = &p->bw;
	const int exact = o->norm != HEM_NORM_TREE;
	const size_t TR = (size_t)p->n_tx * rc.R;
	double *xr = exact ? w->X : NULL, *xv = exact ? w->X + TR : NULL;
	HEM_TRY_DEV(hemk_tile_map12(&p->dt, rc, chunk, w->cnt,
	                            o->precision == HEM_PREC_FP32, w->uq, w->A,
	                            w->B, w->C, exact ? NULL : w->part, xr, xv,
	                            p->n_tx, o->step, p->stream),
	            "tile map12");
	BMARK(HEM_BK_MAP12);
	return HEM_OK;
}

/* One SQUAREM iteration of every active instance, its u1 and u2 already
 * in B and C: the norms, then one kernel that extrapolates, maps with the
 * stop test and maps twice more for the next iteration. */
static int batch_iteration(hem_problem *p, const hem_opts *o, hemk_rctl rc,
                           int chunk, bprof *pc)
{
	hem_batch_ws *w = &p->bw;
	void *s = p->stream;
	/* exact sums by speculation for one sample; several samples (rows of
	 * different lengths) use the single-lane chains, with the same
	 * results */
	const int spec = hem_norm_exact(o->norm) && p->ns == 1;
	const int exact = o->norm != HEM_NORM_TREE;
	/* exact: alpha from the sums inside ex3; tree: alpha given */
	hemk_step st = {o->step, o->clamp, o->alpha_clamp != 0, !exact};
	const size_t TR = (size_t)p->n_tx * rc.R;
	double *xr = exact ? w->X : NULL, *xv = exact ? w->X + TR : NULL;
	if (spec) {
		/* the reference's left-to-right sums by speculation: pass 1
		 * sums (sd) or sums of squares (L2, straight into q_r/q_v),
		 * then the centred sums of the sd form into q_r/q_v */
		const int sd = o->step == HEM_STEP_SD;
		double *sums1 = pass1_sums(p);
		hemk_spec sp = spec_batch(p, o, rc, 1,
		                          sd ? HEMK_SPEC_IDENT : HEMK_SPEC_SQUARE,
		                          NULL, sd ? sums1 : rc.q_r);
		/* with chunk programs, pass 1 keeps each chunk's sum of squares
		 * and its published sum, from which pass 2 (the same chain
		 * slots) seeds its guesses */
		const int seed = sd && sp.cpub;
		if (seed)
			sp.csq = w->scsq;
		HEM_TRY_DEV(hemk_spec_sum(&sp, 0, NULL, NULL, NULL, NULL, s),
		            "batch exact sums 1");
		BMARK(HEM_BK_NORM1);
		if (sd) {
			sp = spec_batch(p, o, rc, 1, HEMK_SPEC_CENTRED, sums1,
			                rc.q_r);
			if (seed) {
				sp.gsum = (const double *)w->scpub;
				sp.gsq = w->scsq;
			}
			HEM_TRY_DEV(hemk_spec_sum(&sp, 0, NULL, NULL, NULL, NULL,
			                          s),
			            "batch exact sums 2");
			BMARK(HEM_BK_NORM2);
		}
	} else if (exact) {
		if (o->step == HEM_STEP_SD) {
			HEM_TRY_DEV(hemk_batch_chain(rc, p->n_tx, xr, xv, 1, s),
			            "batch chain 1");
			BMARK(HEM_BK_NORM1);
			HEM_TRY_DEV(hemk_batch_chain(rc, p->n_tx, xr, xv, 2, s),
			            "batch chain 2");
			BMARK(HEM_BK_NORM2);
		} else {
			HEM_TRY_DEV(hemk_batch_chain(rc, p->n_tx, xr, xv, 3, s),
			            "batch chain");
			BMARK(HEM_BK_NORM1);
		}
	} else {
		HEM_TRY_DEV(hemk_batch_alpha(&p->dt, rc, w->part, p->n_tx, st, s),
		            "batch alpha");
		BMARK(HEM_BK_NORM1);
	}
	HEM_TRY_DEV(hemk_tile_ex3(&p->dt, rc, chunk, w->cnt,
	                          o->precision == HEM_PREC_FP32, w->uq, w->A,
	                          w->B, w->C, p->n_tx, st, 1,
	                          exact ? NULL : w->part, xr, xv, s),
	            "tile ex3");
	BMARK(HEM_BK_EX3);
	HEM_TRY_DEV(hemk_batch_iter_end(rc, s), "batch iter end");
	BMARK(HEM_BK_ITER_END);
	return HEM_OK;
}

static int batch_graph_key(const hem_opts *o, int chunk)
{
	return o->step | (o->clamp << 1) | ((o->alpha_clamp != 0) << 2) |
	       (o->norm << 3) | (chunk << 5) |
	       ((o->precision == HEM_PREC_FP32) << 12);
}

static int batch_graph(hem_problem *p, const hem_opts *o, hemk_rctl rc,
                       int chunk, void **graph)
{
	hem_batch_ws *w = &p->bw;
	const int key = batch_graph_key(o, chunk);
	void *exec = NULL;
	int rc2, e;
	for (int g = 0; g < w->n_graph; ++g) {
		const hem_opts *q = &w->graph_opts[g];
		if (w->graph_key[g] == key && w->graph_R[g] == rc.R &&
		    q->min_iter == o->min_iter && q->max_iter == o->max_iter &&
		    q->stop_min_value == o->stop_min_value &&
		    q->stop_rel_change == o->stop_rel_change) {
			*graph = w->graph[g];
			return HEM_OK;
		}
	}
	if (w->n_graph == HEM_GRAPH_CACHE && (rc2 = ws_retire_graphs(p)))
		return rc2;
	e = hemd_graph_begin(p->stream);
	if (e)
		return hem_dev_fail(e, "batch graph capture");
	for (int i = 0; i < HEM_BATCH_ITERS_PER_GRAPH; ++i) {
		rc2 = batch_iteration(p, o, rc, chunk, NULL);
		if (rc2) {
			hemd_graph_end(p->stream, &exec);
			hemd_graph_destroy(exec);
			return rc2;
		}
	}
	e = hemd_graph_end(p->stream, &exec);
	if (e)
		return hem_dev_fail(e, "batch graph capture end");
	w->graph[w->n_graph] = exec;
	w->graph_key[w->n_graph] = key;
	w->graph_opts[w->n_graph] = *o;
	w->graph_R[w->n_graph++] = rc.R;
	*graph = exec;
	return HEM_OK;
}

static int check_batch_opts(const hem_problem *p, const hem_opts *o)
{
	if (p->weighted)
		return hem_fail(HEM_ERR_INVALID, "batched solves take tables "
		                                 "without member weights");
	if (o->precision != HEM_PREC_FP64 && o->precision != HEM_PREC_FP32)
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
	if (o->fallback)
		return hem_fail(HEM_ERR_INVALID, "batched solves do not take the "
		                                 "monotonicity fallback");
	if (!p->dt.n_tiles)
		return hem_fail(HEM_ERR_RANGE,
		                "a connected component of this class table exceeds "
		                "the tile limits (%u transcripts / %u classes); "
		                "batched solves need tiles", HEM_TILE_TX,
		                HEM_TILE_CLS);
	return HEM_OK;
}

/* Counts already in w->cat_rm ([R][K], caller order): transpose to
 * [K][R] and gather the unique and class rows into tile order. */
static int stage_counts(hem_problem *p, uint32_t R)
{
	hem_batch_ws *w = &p->bw;
	const uint32_t K = p->n_tx + p->n_cls;
	HEM_TRY_DEV(hemk_transpose_u32(w->cat_rm, R, K, w->catT, p->stream),
	            "count transpose");
	HEM_TRY_DEV(hemk_gather_rows_u32(w->catT, p->dt.tx, p->n_tx, R, w->uq,
	                                 p->stream),
	            "unique row gather");
	HEM_TRY_DEV(hemk_gather_rows_u32(w->catT + (size_t)p->n_tx * R,
	                                 p->d_cls_orig, p->dt.n_cls, R, w->cnt,
	                                 p->stream),
	            "class row gather");
	return HEM_OK;
}

/* Runs the loop for R replicates whose counts are staged, then writes the
 * outputs. */
static int run_batch(hem_problem *p, const hem_opts *o, uint32_t R,
                     double *est, double *tpm, uint32_t *iters,
                     hem_batch_stats *st)
{
	hem_batch_ws *w = &p->bw;
	hemk_rctl rc = rctl_for(p, R, o);
	const int chunk = hem_batch_chunk(p, R);
	void *graph = NULL;
	uint32_t launched = 0;
	const uint32_t limit = o->max_iter > o->min_iter ? o->max_iter
	                                                  : o->min_iter;
	float ms = 0;
	int rc2;

	if (chunk <= 0)
		return hem_fail(HEM_ERR_RANGE, "tiles do not fit the LDS budget");
	rc.chunk = (uint32_t)chunk;
	HEM_TRY_DEV(hemk_batch_reset(rc, p->n_tx, w->A, p->stream), "reset");
	HEM_TRY_DEV(hemk_fill_u32(reruns_word(p), 1, 0, p->stream),
	            "rerun counter");
	if ((rc2 = batch_prologue(p, o, rc, chunk, NULL)))
		return rc2;
	if ((rc2 = batch_graph(p, o, rc, chunk, &graph)))
		return rc2;
	HEM_TRY_DEV(hemd_event_record(p->ev0, p->stream), "event");
	for (;;) {
		HEM_TRY_DEV(hemd_graph_launch(graph, p->stream), "graph launch");
		launched += HEM_BATCH_ITERS_PER_GRAPH;
		if (launched < o->min_iter && launched < limit)
			continue;
		HEM_TRY_DEV(hemd_d2h(w->h_active, rc.active,
		                     (size_t)p->ns * 4, p->stream),
		            "active readback");
		HEM_TRY_DEV(hemd_stream_sync(p->stream), "stream sync");
		{
			uint32_t act = 0;
			for (uint32_t q = 0; q < p->ns; ++q)
				act |= w->h_active[q];
			if (act == 0)
				break;
		}
		if (launched > limit + HEM_BATCH_ITERS_PER_GRAPH)
			return hem_fail(HEM_ERR_DEVICE,
			                "batched loop did not terminate");
	}
	HEM_TRY_DEV(hemd_event_record(p->ev1, p->stream), "event");
	/* finalisation: the state as one row per replicate in transcript
	 * order, its sequential sum (src/EM.c:381-384), then est/tpm [r][t]
	 * into B and C */
	HEM_TRY_DEV(hemk_batch_rows(rc, p->n_tx, p->dt.txq, w->A, w->X,
	                            p->stream),
	            "state rows");
	if (o->norm == HEM_NORM_SEQ || p->ns > 1) {
		HEM_TRY_DEV(hemk_batch_chain(rc, p->n_tx, w->X, NULL, 0,
		                             p->stream),
		            "final sums");
	} else {
		hemk_spec sp = spec_batch(p, o, rc, 0, HEMK_SPEC_IDENT, NULL,
		                          rc.lam_sum);
		HEM_TRY_DEV(hemk_spec_sum(&sp, 0, NULL, NULL, NULL, NULL,
		                          p->stream),
		            "final sums");
	}
	HEM_TRY_DEV(hemk_batch_out(rc, p->n_tx, w->X, p->d_len, w->B, w->C,
	                           p->stream),
	            "finalize");
	if (est)
		HEM_TRY_DEV(hemd_d2h(est, w->B, (size_t)p->n_tx * R * 8,
		                     p->stream),
		            "est readback");
	if (tpm)
		HEM_TRY_DEV(hemd_d2h(tpm, w->C, (size_t)p->n_tx * R * 8,
		                     p->stream),
		            "tpm readback");
	{
		const size_t NI = (size_t)p->ns * R;
		uint32_t *hk = malloc(NI * 4 * 2 + 4);
		if (!hk)
			return hem_fail(HEM_ERR_NOMEM, "host allocation");
		rc2 = hemd_d2h(hk, rc.k, NI * 4, p->stream);
		if (!rc2)
			rc2 = hemd_d2h(hk + NI, rc.conv, NI * 4, p->stream);
		if (!rc2)
			rc2 = hemd_d2h(hk + 2 * NI, reruns_word(p), 4, p->stream);
		if (!rc2)
			rc2 = hemd_stream_sync(p->stream);
		if (rc2) {
			free(hk);
			return hem_dev_fail(rc2, "iteration readback");
		}
		if (iters)
			memcpy(iters, hk, NI * 4);
		if (st) {
			st->n_rep = R;
			st->iters_min = 0xFFFFFFFFu;
			st->iters_max = 0;
			st->n_converged = 0;
			for (size_t r = 0; r < NI; ++r) {
				if (hk[r] < st->iters_min)
					st->iters_min = hk[r];
				if (hk[r] > st->iters_max)
					st->iters_max = hk[r];
				st->n_converged += hk[NI + r] != 0;
			}
			st->chunk = (uint32_t)chunk;
			st->sum_reruns = hk[2 * NI];
			st->n_tiles = p->dt.n_tiles;
		}
		free(hk);
	}
	HEM_TRY_DEV(hemd_event_elapsed(p->ev0, p->ev1, &ms), "event time");
	if (st)
		st->ms_solve = ms;
	return HEM_OK;
}

static int load_batch_opts(const hem_opts *opts, hem_opts *o)
{
	if (opts)
		*o = *opts;
	else
		hem_opts_default(o);
	return HEM_OK;
}

/* The counts of n_rep solves as rows [r][T + C] (unique counts, then class
 * counts as u32), checked; NULL with the error set on failure. */
static uint32_t *pack_counts(const hem_problem *p, uint32_t n_rep,
                             const uint32_t *unique, const double *cls_count)
{
	const uint32_t T = p->n_tx, C = p->n_cls;
	const size_t K = (size_t)T + C;
	uint32_t *h;
	if (!cls_count && !p->counts_integral) {
		hem_fail(HEM_ERR_INVALID,
		         "class counts must be integers below 2^32");
		return NULL;
	}
	h = malloc(K * n_rep * 4);
	if (!h) {
		hem_fail(HEM_ERR_NOMEM, "host allocation");
		return NULL;
	}
	for (uint32_t r = 0; r < n_rep; ++r) {
		uint32_t *row = h + (size_t)r * K;
		memcpy(row, unique + (size_t)r * T, (size_t)T * 4);
		for (uint32_t c = 0; c < C; ++c) {
			const double x = cls_count ? cls_count[(size_t)r * C + c]
			                           : p->h_cls_count[c];
			if (!(x >= 0) || x > 4294967295.0 || x != floor(x)) {
				free(h);
				hem_fail(HEM_ERR_INVALID,
				         "class count %g (replicate %u, class "
				         "%u) is not an integer below 2^32",
				         x, r, c);
				return NULL;
			}
			row[T + c] = (uint32_t)x;
		}
	}
	return h;
}

/* Uploads packed counts into the workspace (sized for n_rep) and stages
 * them in tile order; p->mu held. */
static int upload_counts(hem_problem *p, uint32_t n_rep, const uint32_t *h)
{
	const size_t K = (size_t)p->n_tx + p->n_cls;
	int rc;
	if ((rc = ws_ensure(p, n_rep)))
		return rc;
	rc = hemd_h2d(p->bw.cat_rm, h, K * n_rep * 4, p->stream);
	if (rc)
		return hem_dev_fail(rc, "count upload");
	return stage_counts(p, n_rep);
}

int hem_solve_batch(hem_problem *p, uint32_t n_rep, const uint32_t *unique,
                    const double *cls_count, const hem_opts *opts, double *est,
                    double *tpm, uint32_t *iters, hem_batch_stats *st)
{
	hem_opts o;
	int rc;
	const double t0 = now_ms();
	uint32_t *h = NULL;
	if (!p || !unique || n_rep == 0)
		return hem_fail(HEM_ERR_INVALID, "null problem/counts or n_rep 0");
	load_batch_opts(opts, &o);
	if ((rc = check_batch_opts(p, &o)))
		return rc;
	if (!(h = pack_counts(p, n_rep, unique, cls_count)))
		return hem_last_code();
	pthread_mutex_lock(&p->mu);
	if ((rc = upload_counts(p, n_rep, h)))
		goto out;
	rc = run_batch(p, &o, n_rep, est, tpm, iters, st);
	if (!rc && st)
		st->ms_total = now_ms() - t0;
out:
	pthread_mutex_unlock(&p->mu);
	free(h);
	return rc;
}

/* Algorithmic bytes of one launch over `act` active replicates, with u32
 * indices and FP32 values: per EM map the structure (8Z + 4C + 8T) once
 * and each replicate's counts, densities and outputs (12C + 16T); each
 * exact-sum pass reads the r and v rows of every active replicate. */
static double batch_launch_bytes(const hem_problem *p, int kind, double act)
{
	const double T = p->n_tx, C = p->n_cls, Z = (double)p->nnz;
	const double shared = 8 * Z + 4 * C + 8 * T, per = 12 * C + 16 * T;
	switch (kind) {
	case HEM_BK_MAP12:
		return 2 * (shared + act * per);
	case HEM_BK_EX3:
		return 3 * (shared + act * per);
	case HEM_BK_NORM1:
	case HEM_BK_NORM2:
		return 16 * T * act;
	default:
		return 8 * act;
	}
}

int hem_profile_batch(hem_problem *p, uint32_t n_rep, const uint32_t *unique,
                      const double *cls_count, const hem_opts *opts,
                      hem_batch_profile *out)
{
	hem_opts o;
	bprof pc;
	void *ev_start = NULL;
	uint32_t *h = NULL, act_h[1];
	int rc = HEM_OK, i, e;
	if (!p || !unique || n_rep == 0 || !out)
		return hem_fail(HEM_ERR_INVALID, "null argument or n_rep 0");
	load_batch_opts(opts, &o);
	if ((rc = check_batch_opts(p, &o)))
		return rc;
	if (!(h = pack_counts(p, n_rep, unique, cls_count)))
		return hem_last_code();
	memset(out, 0, sizeof(*out));
	memset(&pc, 0, sizeof(pc));
	for (i = 0; i < 16 && !rc; ++i)
		rc = hemd_event_create(&pc.ev[i]);
	if (!rc)
		rc = hemd_event_create(&ev_start);
	if (rc) {
		rc = hem_dev_fail(rc, "event creation");
		goto free_ev;
	}
	pthread_mutex_lock(&p->mu);
	if ((rc = upload_counts(p, n_rep, h)))
		goto out;
	{
		hem_batch_ws *w = &p->bw;
		hemk_rctl bc = rctl_for(p, n_rep, &o);
		const int chunk = hem_batch_chunk(p, n_rep);
		const uint32_t limit = o.max_iter > o.min_iter ? o.max_iter
		                                                : o.min_iter;
		uint32_t act = n_rep * p->ns;
		if (chunk <= 0) {
			rc = hem_fail(HEM_ERR_RANGE,
			              "tiles do not fit the LDS budget");
			goto out;
		}
		bc.chunk = (uint32_t)chunk;
		e = hemk_batch_reset(bc, p->n_tx, w->A, p->stream);
		if (!e)
			e = hemk_fill_u32(reruns_word(p), 1, 0, p->stream);
		if (e) {
			rc = hem_dev_fail(e, "reset");
			goto out;
		}
		for (int it = -1;; ++it) {
			pc.n = 0;
			e = hemd_event_record(ev_start, p->stream);
			if (e) {
				rc = hem_dev_fail(e, "event");
				goto out;
			}
			rc = it < 0 ? batch_prologue(p, &o, bc, chunk, &pc)
			            : batch_iteration(p, &o, bc, chunk, &pc);
			if (rc)
				goto out;
			e = hemd_d2h(w->h_active, bc.active,
			             (size_t)p->ns * 4, p->stream);
			if (!e)
				e = hemd_stream_sync(p->stream);
			if (e) {
				rc = hem_dev_fail(e, "profiled iteration");
				goto out;
			}
			for (i = 0; i < pc.n; ++i) {
				float ms = 0;
				e = hemd_event_elapsed(i ? pc.ev[i - 1] : ev_start,
				                       pc.ev[i], &ms);
				if (e) {
					rc = hem_dev_fail(e, "event time");
					goto out;
				}
				out->ms[pc.kind[i]] += ms;
				out->launches[pc.kind[i]]++;
				out->bytes[pc.kind[i]] +=
				    batch_launch_bytes(p, pc.kind[i], act);
			}
			if (it < 0)
				continue;
			out->iters++;
			out->rep_iters += act;
			act = 0;
			for (uint32_t q = 0; q < p->ns; ++q)
				act += w->h_active[q];
			if (act == 0)
				break;
			if (out->iters > limit + 1) {
				rc = hem_fail(HEM_ERR_DEVICE,
				              "profiled loop did not terminate");
				goto out;
			}
		}
		out->chunk = (uint32_t)chunk;
		out->n_tiles = p->dt.n_tiles;
		{
			/* the exact sums' reruns, as run_batch reports them */
			e = hemd_d2h(act_h, reruns_word(p), 4, p->stream);
			if (!e)
				e = hemd_stream_sync(p->stream);
			if (e) {
				rc = hem_dev_fail(e, "rerun readback");
				goto out;
			}
			out->sum_reruns = act_h[0];
		}
	}
out:
	pthread_mutex_unlock(&p->mu);
free_ev:
	for (i = 0; i < 16; ++i)
		if (pc.ev[i])
			hemd_event_destroy(pc.ev[i]);
	if (ev_start)
		hemd_event_destroy(ev_start);
	free(h);
	return rc;
}

/* ---- resampling ---------------------------------------------------------- */

/* Splits K weighted categories into chunks: a category heavier than the
 * target weight W is a chunk on its own; otherwise categories join the open
 * chunk until it would pass W or hold HEMK_RS_CHUNK categories.  W starts at
 * ceil(N / 2048) and doubles until the chunk count fits HEMK_RS_MAXB. */
static int plan_chunks(const uint32_t *w, uint32_t K, uint32_t N,
                       uint32_t *start, uint32_t *cw, uint32_t *B_out)
{
	uint64_t W = ((uint64_t)N + 2047) / 2048;
	if (W == 0)
		W = 1;
	for (;;) {
		uint32_t B = 0, n = 0;
		uint64_t cur = 0, acc = 0;
		int over = 0;
		for (uint32_t k = 0; k < K && !over; ++k) {
			if (w[k] > W) {
				if (n) {
					cw[B] = (uint32_t)acc;
					start[B++] = k - n;
					acc += cur;
					n = 0;
					cur = 0;
				}
				if (B >= HEMK_RS_MAXB) {
					over = 1;
					break;
				}
				cw[B] = (uint32_t)acc;
				start[B++] = k;
				acc += w[k];
				continue;
			}
			if (n && (cur + w[k] > W || n == HEMK_RS_CHUNK)) {
				cw[B] = (uint32_t)acc;
				start[B++] = k - n;
				acc += cur;
				n = 0;
				cur = 0;
			}
			if (B >= HEMK_RS_MAXB) {
				over = 1;
				break;
			}
			cur += w[k];
			++n;
		}
		if (!over && n) {
			if (B >= HEMK_RS_MAXB) {
				over = 1;
			} else {
				cw[B] = (uint32_t)acc;
				start[B++] = K - n;
				acc += cur;
			}
		}
		if (!over) {
			start[B] = K;
			cw[B] = (uint32_t)acc;
			*B_out = B;
			return HEM_OK;
		}
		W *= 2;
	}
}

/* Resamples n_rep count vectors of every sample into w->cat_rm ([R][K],
 * K = n_tx + n_cls of the handle: unique counts, then class counts);
 * sample q uses seeds[q * n_rep + r]. */
static int resample_one(hem_problem *p, const uint32_t *unique,
                        uint32_t n_rep, const uint64_t *seeds, uint32_t q,
                        float *ms);

static int resample(hem_problem *p, const uint32_t *unique, uint32_t n_rep,
                    const uint64_t *seeds, float *ms)
{
	float tot = 0;
	for (uint32_t q = 0; q < p->ns; ++q) {
		float m = 0;
		int rc = resample_one(p, unique, n_rep, seeds + (size_t)q * n_rep,
		                      q, ms ? &m : NULL);
		if (rc)
			return rc;
		tot += m;
	}
	if (ms)
		*ms = tot;
	return HEM_OK;
}

static int resample_one(hem_problem *p, const uint32_t *unique,
                        uint32_t n_rep, const uint64_t *seeds, uint32_t q,
                        float *ms)
{
	hem_batch_ws *wk = &p->bw;
	const uint32_t t0 = p->ns > 1 ? p->h_tx0[q] : 0;
	const uint32_t c0 = p->ns > 1 ? p->h_cls0[q] : 0;
	const uint32_t T = p->ns > 1 ? p->h_tx0[q + 1] - t0 : p->n_tx;
	const uint32_t C = p->ns > 1 ? p->h_cls0[q + 1] - c0 : p->n_cls;
	const uint32_t K = T + C;
	uint32_t *hw = malloc(((size_t)K + 2 * (HEMK_RS_MAXB + 1)) * 4);
	uint32_t *start = hw + K, *cw = start + HEMK_RS_MAXB + 1, B = 0;
	uint32_t *d = NULL;
	uint64_t N = 0;
	int rc;
	if (!hw)
		return hem_fail(HEM_ERR_NOMEM, "host allocation");
	if (!p->counts_integral) {
		free(hw);
		return hem_fail(HEM_ERR_INVALID,
		                "bootstrap needs integer class counts below 2^32");
	}
	for (uint32_t t = 0; t < T; ++t) {
		hw[t] = unique[t0 + t];
		N += unique[t0 + t];
	}
	for (uint32_t c = 0; c < C; ++c) {
		hw[T + c] = (uint32_t)p->h_cls_count[c0 + c];
		N += hw[T + c];
	}
	if (N > 0xFFFFFFFFull) {
		free(hw);
		return hem_fail(HEM_ERR_RANGE,
		                "%llu fragments: the bootstrap needs fewer than "
		                "2^32", (unsigned long long)N);
	}
	plan_chunks(hw, K, (uint32_t)N, start, cw, &B);
	d = wk->rs_plan;
	rc = hemd_h2d(d, hw, ((size_t)K + 2 * (HEMK_RS_MAXB + 1)) * 4,
	              p->stream);
	if (!rc)
		rc = hemd_h2d(wk->seeds, seeds, (size_t)n_rep * 8, p->stream);
	if (!rc)
		rc = hemd_event_record(p->ev2, p->stream);
	if (!rc)
		rc = hemk_resample(d, K, d + K, d + K + HEMK_RS_MAXB + 1, B,
		                   (uint32_t)N, wk->seeds, n_rep, wk->m,
		                   p->ns > 1 ? wk->catT : wk->cat_rm, p->stream);
	/* several samples: this sample's rows into its ranges of the
	 * handle's unique and class blocks */
	if (!rc && p->ns > 1)
		rc = hemk_scatter_counts(wk->catT, n_rep, T, C, wk->cat_rm,
		                         p->n_tx + p->n_cls, t0, p->n_tx + c0,
		                         p->stream);
	if (!rc)
		rc = hemd_event_record(p->ev3, p->stream);
	if (!rc)
		rc = hemd_event_sync(p->ev3);
	if (!rc && ms)
		rc = hemd_event_elapsed(p->ev2, p->ev3, ms);
	free(hw);
	return rc ? hem_dev_fail(rc, "resampling") : HEM_OK;
}

int hem_bootstrap(hem_problem *p, const uint32_t *unique, uint32_t n_rep,
                  const uint64_t *seeds, const hem_opts *opts, double *est,
                  double *tpm, uint32_t *iters, hem_batch_stats *st)
{
	hem_opts o;
	int rc;
	float ms_rs = 0;
	const double t0 = now_ms();
	if (!p || !unique || !seeds || n_rep == 0)
		return hem_fail(HEM_ERR_INVALID, "null argument or n_rep 0");
	load_batch_opts(opts, &o);
	if ((rc = check_batch_opts(p, &o)))
		return rc;
	pthread_mutex_lock(&p->mu);
	if ((rc = ws_ensure(p, n_rep)))
		goto out;
	if ((rc = resample(p, unique, n_rep, seeds, &ms_rs)))
		goto out;
	if ((rc = stage_counts(p, n_rep)))
		goto out;
	rc = run_batch(p, &o, n_rep, est, tpm, iters, st);
	if (!rc && st) {
		st->ms_resample = ms_rs;
		st->ms_total = now_ms() - t0;
	}
out:
	pthread_mutex_unlock(&p->mu);
	return rc;
}

/* hera's sampler into w->cat_rm (one-sample handles): unique counts
 * redrawn, class counts the table's (src/EM.c:98-150). */
static int resample_hera(hem_problem *p, const uint32_t *unique,
                         uint32_t n_rep, uint32_t seed, float *ms)
{
	const uint32_t T = p->n_tx, C = p->n_cls;
	uint32_t *h = NULL, *d = NULL, sh = 0;
	uint64_t N = 0;
	int rc = HEM_OK, e = 0;
	if (p->ns > 1)
		return hem_fail(HEM_ERR_INVALID, "hera's sampler is for "
		                                 "one-sample handles");
	if (!p->counts_integral)
		return hem_fail(HEM_ERR_INVALID,
		                "bootstrap needs integer class counts below 2^32");
	if (n_rep > 65535)
		return hem_fail(HEM_ERR_RANGE, "at most 65535 replicates");
	h = malloc(((size_t)T + C) * 4);
	if (!h)
		return hem_fail(HEM_ERR_NOMEM, "host allocation");
	for (uint32_t t = 0; t < T; ++t) {
		N += unique[t];
		h[t] = (uint32_t)N;
		if (N > 0xFFFFFFFFull) {
			free(h);
			return hem_fail(HEM_ERR_RANGE, "%llu unique reads: hera's "
			                "sampler needs fewer than 2^32",
			                (unsigned long long)N);
		}
	}
	for (uint32_t c = 0; c < C; ++c)
		h[T + c] = (uint32_t)p->h_cls_count[c];
	/* about four slot buckets per transcript */
	while (N > 1 && ((N - 1) >> sh) + 1 > 4 * (uint64_t)T)
		++sh;
	{
		const size_t nb = N > 1 ? (size_t)(((N - 1) >> sh) + 1) : 1;
		e = hemd_malloc((void **)&d, ((size_t)T + C + nb) * 4);
		if (e) {
			free(h);
			return hem_fail(HEM_ERR_NOMEM, "device allocation: %s",
			                hemd_error_string(e));
		}
		e = hemd_h2d(d, h, ((size_t)T + C) * 4, p->stream);
		if (!e)
			e = hemd_event_record(p->ev2, p->stream);
		if (!e)
			e = hemk_hera_resample(d, T, (uint32_t)N, d + T + C, sh,
			                       d + T, C, seed, n_rep, p->bw.cat_rm,
			                       p->stream);
		if (!e)
			e = hemd_event_record(p->ev3, p->stream);
		if (!e)
			e = hemd_event_sync(p->ev3);
		if (!e && ms)
			e = hemd_event_elapsed(p->ev2, p->ev3, ms);
	}
	if (e)
		rc = hem_dev_fail(e, "hera's sampler");
	hemd_free(d);
	free(h);
	return rc;
}

int hem_bootstrap_hera(hem_problem *p, const uint32_t *unique,
                       uint32_t n_rep, uint32_t seed, const hem_opts *opts,
                       double *est, double *tpm, uint32_t *iters,
                       hem_batch_stats *st)
{
	hem_opts o;
	int rc;
	float ms_rs = 0;
	const double t0 = now_ms();
	if (!p || !unique || n_rep == 0)
		return hem_fail(HEM_ERR_INVALID, "null argument or n_rep 0");
	load_batch_opts(opts, &o);
	if ((rc = check_batch_opts(p, &o)))
		return rc;
	pthread_mutex_lock(&p->mu);
	if ((rc = ws_ensure(p, n_rep)))
		goto out;
	if ((rc = resample_hera(p, unique, n_rep, seed, &ms_rs)))
		goto out;
	if ((rc = stage_counts(p, n_rep)))
		goto out;
	rc = run_batch(p, &o, n_rep, est, tpm, iters, st);
	if (!rc && st) {
		st->ms_resample = ms_rs;
		st->ms_total = now_ms() - t0;
	}
out:
	pthread_mutex_unlock(&p->mu);
	return rc;
}

int hem_bootstrap_hera_counts(hem_problem *p, const uint32_t *unique,
                              uint32_t n_rep, uint32_t seed,
                              uint32_t *unique_out)
{
	int rc;
	const uint32_t T = p ? p->n_tx : 0, C = p ? p->n_cls : 0;
	const size_t K = (size_t)T + C;
	uint32_t *h;
	if (!p || !unique || !unique_out || n_rep == 0)
		return hem_fail(HEM_ERR_INVALID, "null argument or n_rep 0");
	h = malloc(K * n_rep * 4);
	if (!h)
		return hem_fail(HEM_ERR_NOMEM, "host allocation");
	pthread_mutex_lock(&p->mu);
	rc = ws_ensure(p, n_rep);
	if (!rc)
		rc = resample_hera(p, unique, n_rep, seed, NULL);
	if (!rc) {
		int e = hemd_d2h(h, p->bw.cat_rm, K * n_rep * 4, p->stream);
		if (!e)
			e = hemd_stream_sync(p->stream);
		if (e)
			rc = hem_dev_fail(e, "count readback");
	}
	pthread_mutex_unlock(&p->mu);
	if (!rc)
		for (uint32_t r = 0; r < n_rep; ++r)
			memcpy(unique_out + (size_t)r * T, h + (size_t)r * K,
			       (size_t)T * 4);
	free(h);
	return rc;
}

int hem_bootstrap_counts(hem_problem *p, const uint32_t *unique,
                         uint32_t n_rep, const uint64_t *seeds,
                         uint32_t *unique_out, double *cls_out)
{
	int rc;
	const uint32_t T = p ? p->n_tx : 0, C = p ? p->n_cls : 0;
	const size_t K = (size_t)T + C;
	uint32_t *h;
	if (!p || !unique || !seeds || n_rep == 0)
		return hem_fail(HEM_ERR_INVALID, "null argument or n_rep 0");
	h = malloc(K * n_rep * 4);
	if (!h)
		return hem_fail(HEM_ERR_NOMEM, "host allocation");
	pthread_mutex_lock(&p->mu);
	rc = ws_ensure(p, n_rep);
	if (!rc)
		rc = resample(p, unique, n_rep, seeds, NULL);
	if (!rc) {
		int e = hemd_d2h(h, p->bw.cat_rm, K * n_rep * 4, p->stream);
		if (!e)
			e = hemd_stream_sync(p->stream);
		if (e)
			rc = hem_dev_fail(e, "count readback");
	}
	pthread_mutex_unlock(&p->mu);
	if (!rc)
		for (uint32_t r = 0; r < n_rep; ++r) {
			const uint32_t *row = h + (size_t)r * K;
			if (unique_out)
				memcpy(unique_out + (size_t)r * T, row, (size_t)T * 4);
			if (cls_out)
				for (uint32_t c = 0; c < C; ++c)
					cls_out[(size_t)r * C + c] = row[T + c];
		}
	free(h);
	return rc;
}
