/*
 * tune.c -- the tile configuration autotuner and its cache file.
 *
 * No counterpart in the reference (its EM is a CPU loop).  The tile kernels'
 * speed depends on how components are packed into tiles and, in the batched
 * kernels, on how many replicates share a workgroup; neither changes a
 * result in the order-faithful modes, so the tuner only has to time.  Each
 * candidate is a temporary handle (tilings are built when a handle is
 * created); the replicate chunk of the batched kernels is varied on a live
 * handle, since it is read at launch.
 *
 * Cache file: text, one configuration per line,
 *   <key> ptile_tx ptile_cls ptile_nnz tile_tx tile_cls tile_nnz chunk
 * with key "T=<n_tx>,C=<n_cls>,Z=<members>,R=<n_rep>,arch=<device arch>";
 * lines starting with '#' are comments; the last line of a key wins.
 */
#include <errno.h>
#include <float.h>
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hem_internal.h"

int hem_tune_key(uint32_t n_tx, uint32_t n_cls, uint64_t nnz, uint32_t n_rep,
                 const char *arch, char *key, size_t len)
{
	int n;
	if (!arch || !key || len == 0)
		return hem_fail(HEM_ERR_INVALID, "null argument");
	for (const char *c = arch; *c; ++c)
		if (*c == ' ' || *c == '\t' || *c == '\n')
			return hem_fail(HEM_ERR_INVALID,
			                "arch name holds white space");
	n = snprintf(key, len, "T=%u,C=%u,Z=%" PRIu64 ",R=%u,arch=%s", n_tx,
	             n_cls, nnz, n_rep, arch);
	if (n < 0 || (size_t)n >= len)
		return hem_fail(HEM_ERR_RANGE, "key buffer too small");
	return HEM_OK;
}

int hem_tune_cache_lookup(const char *path, const char *key,
                          hem_tile_config *cfg)
{
	char line[512], k[256];
	hem_tile_config g, found;
	int have = 0;
	FILE *f;
	if (!path || !key || !cfg)
		return hem_fail(HEM_ERR_INVALID, "null argument");
	f = fopen(path, "r");
	if (!f)
		return hem_fail(HEM_ERR_INVALID, "tuning cache %s: %s", path,
		                strerror(errno));
	while (fgets(line, sizeof(line), f)) {
		if (line[0] == '#' || line[0] == '\n')
			continue;
		if (sscanf(line, "%255s %u %u %u %u %u %u %u", k, &g.ptile_tx,
		           &g.ptile_cls, &g.ptile_nnz, &g.tile_tx, &g.tile_cls,
		           &g.tile_nnz, &g.chunk) != 8)
			continue;
		if (strcmp(k, key) == 0) {
			found = g;
			have = 1;
		}
	}
	fclose(f);
	if (!have)
		return hem_fail(HEM_ERR_INVALID, "no configuration for %s in %s",
		                key, path);
	*cfg = found;
	return HEM_OK;
}

int hem_tune_cache_store(const char *path, const char *key,
                         const hem_tile_config *g)
{
	FILE *f;
	int bad;
	if (!path || !key || !g)
		return hem_fail(HEM_ERR_INVALID, "null argument");
	if (strchr(key, ' ') || strchr(key, '\n'))
		return hem_fail(HEM_ERR_INVALID, "key holds white space");
	f = fopen(path, "a");
	if (!f)
		return hem_fail(HEM_ERR_INVALID, "tuning cache %s: %s", path,
		                strerror(errno));
	bad = fprintf(f, "%s %u %u %u %u %u %u %u\n", key, g->ptile_tx,
	              g->ptile_cls, g->ptile_nnz, g->tile_tx, g->tile_cls,
	              g->tile_nnz, g->chunk) < 0;
	bad |= fclose(f) != 0;
	if (bad)
		return hem_fail(HEM_ERR_INVALID, "tuning cache %s: write failed",
		                path);
	return HEM_OK;
}

/* ---- timing ------------------------------------------------------------- */

typedef struct {
	const uint32_t *unique;   /* n_tx                                    */
	uint32_t *urep;           /* R copies of unique (batched timing)    */
	double *o1, *o2, *o3;     /* outputs, R * n_tx                      */
	uint32_t *it;             /* R                                      */
	hem_opts opts;
	uint32_t repeats;
} tune_ctx;

/* Device time per SQUAREM iteration of a plain exact solve (us): the
 * fastest of `repeats` runs after one that captures the graphs. */
static int time_plain(hem_problem *p, tune_ctx *t, double *us)
{
	hem_stats st;
	double best = DBL_MAX;
	int rc;
	for (uint32_t r = 0; r <= t->repeats; ++r) {
		rc = hem_solve(p, t->unique, &t->opts, t->o1, t->o2, t->o3, &st);
		if (rc)
			return rc;
		if (r > 0 && st.iters && st.ms_solve * 1e3 / st.iters < best)
			best = st.ms_solve * 1e3 / st.iters;
	}
	*us = best;
	return HEM_OK;
}

/* Device time per replicate-iteration of an exact batch of R replicates
 * of the table's own counts (us). */
static int time_batch(hem_problem *p, tune_ctx *t, uint32_t R, double *us)
{
	hem_batch_stats st;
	double best = DBL_MAX;
	int rc;
	for (uint32_t r = 0; r <= t->repeats; ++r) {
		rc = hem_solve_batch(p, R, t->urep, NULL, &t->opts, t->o1, t->o2,
		                     t->it, &st);
		if (rc)
			return rc;
		uint64_t ri = 0;
		for (uint32_t q = 0; q < R; ++q)
			ri += t->it[q];
		if (r > 0 && ri && st.ms_solve * 1e3 / (double)ri < best)
			best = st.ms_solve * 1e3 / (double)ri;
	}
	*us = best;
	return HEM_OK;
}

#define NCAND(a) (sizeof(a) / sizeof((a)[0]))
/* the fraction by which a candidate must beat the default */
#define HEM_TUNE_MARGIN 0.02

/* Single-problem tile targets tried (transcripts, classes, members); the
 * first is the default. */
static const uint32_t plain_cand[][3] = {
    {HEM_PTILE_TX_TARGET, HEM_PTILE_CLS_TARGET, HEM_PTILE_NNZ_TARGET},
    {64, 320, 768},  {64, 320, 1536}, {64, 320, 2048},
    {48, 240, 768},  {32, 160, 512},
};
/* Batched tile targets tried (the first is the default) and chunks. */
static const uint32_t batch_cand[][3] = {
    {HEM_TILE_TX_TARGET, HEM_TILE_CLS_TARGET, HEM_TILE_NNZ_TARGET},
    {16, 80, 384},
    {32, 160, 768},
};
/* Replicate chunks tried: the default (the largest that fits), then each
 * explicit chunk that differs from it and fits. */
static const uint32_t chunk_cand[] = {0, 8, 16, 32};

/* Times the batched chunks on a handle, keeping the fastest so far in
 * *best_us and *best_chunk (*improved: this handle gave it). */
static int time_chunks(hem_problem *p, tune_ctx *t, uint32_t R,
                       hem_tune_report *r, double *best_us,
                       uint32_t *best_chunk, int *improved,
                       double *default_us)
{
	const int dflt = hemk_tile_chunk(&p->dt, R, 4);
	int rc = HEM_OK;
	*improved = 0;
	for (size_t c = 0; !rc && c < NCAND(chunk_cand); ++c) {
		double us;
		p->cfg.chunk = chunk_cand[c];
		if (chunk_cand[c] && (hem_batch_chunk(p, R) != (int)chunk_cand[c] ||
		                      (int)chunk_cand[c] == dflt))
			continue;
		rc = time_batch(p, t, R, &us);
		r->n_candidates++;
		if (!rc && c == 0 && default_us)
			*default_us = us;
		if (!rc && us < *best_us) {
			*best_us = us;
			*best_chunk = chunk_cand[c];
			*improved = 1;
		}
	}
	return rc;
}

int hem_autotune(const hem_classes *cls, const uint32_t *unique,
                 const hem_tune_opts *o, hem_tile_config *cfg,
                 hem_tune_report *rep)
{
	hem_tune_report r;
	hem_tile_config best, g;
	tune_ctx t;
	hem_problem *p = NULL;
	uint32_t R, T;
	uint64_t nnz;
	double us, best_plain = DBL_MAX, best_batch = DBL_MAX;
	int rc;

	if (!cls || !unique || !o || !cfg)
		return hem_fail(HEM_ERR_INVALID, "null argument");
	if ((rc = hem_device_ready()))
		return rc;
	memset(&r, 0, sizeof(r));
	memset(&t, 0, sizeof(t));
	R = o->n_rep;
	T = cls->n_tx;
	nnz = cls->n_cls ? cls->cls_off[cls->n_cls] : 0;
	if ((rc = hem_tune_key(T, cls->n_cls, nnz, R, hem_device_arch(), r.key,
	                       sizeof(r.key))))
		return rc;
	if (o->cache_path &&
	    hem_tune_cache_lookup(o->cache_path, r.key, &g) == HEM_OK) {
		*cfg = g;
		r.from_cache = 1;
		if (rep)
			*rep = r;
		return HEM_OK;
	}

	hem_opts_default(&t.opts);
	t.opts.norm = HEM_NORM_EXACT;
	t.opts.precision = HEM_PREC_FP64;
	t.opts.min_iter = t.opts.max_iter = o->iters ? o->iters : 20;
	t.repeats = o->repeats ? o->repeats : 3;
	t.unique = unique;
	{
		const size_t n = (size_t)(R ? R : 1) * T;
		t.o1 = malloc((n ? n : 1) * sizeof(double));
		t.o2 = malloc((n ? n : 1) * sizeof(double));
		t.o3 = malloc(((size_t)T ? T : 1) * sizeof(double));
		t.it = malloc((size_t)(R ? R : 1) * sizeof(uint32_t));
		t.urep = R ? malloc((n ? n : 1) * sizeof(uint32_t)) : NULL;
		if (!t.o1 || !t.o2 || !t.o3 || !t.it || (R && !t.urep)) {
			rc = hem_fail(HEM_ERR_NOMEM, "host allocation");
			goto out;
		}
		for (uint32_t q = 0; q < R; ++q)
			memcpy(t.urep + (size_t)q * T, unique, (size_t)T * 4);
	}
	hem_tile_config_default(&best);

	/* single-problem tiles */
	for (size_t i = 0; i < NCAND(plain_cand); ++i) {
		hem_tile_config_default(&g);
		g.ptile_tx = plain_cand[i][0];
		g.ptile_cls = plain_cand[i][1];
		g.ptile_nnz = plain_cand[i][2];
		if ((rc = hem_problem_create_config(cls, &g, &p)))
			goto out;
		if (i == 0 && !p->dtp.n_tiles) {
			/* the global passes: the tiles are not used */
			hem_problem_destroy(p);
			p = NULL;
			break;
		}
		rc = time_plain(p, &t, &us);
		r.n_candidates++;
		if (!rc && i == 0)
			r.us_plain_default = us;
		if (!rc && us < best_plain) {
			best_plain = us;
			best.ptile_tx = g.ptile_tx;
			best.ptile_cls = g.ptile_cls;
			best.ptile_nnz = g.ptile_nnz;
		}
		/* batched chunks on the default handle's small tiles */
		if (!rc && i == 0 && R && p->dt.n_tiles && !p->weighted) {
			int imp;
			rc = time_chunks(p, &t, R, &r, &best_batch, &best.chunk,
			                 &imp, &r.us_batch_default);
		}
		hem_problem_destroy(p);
		p = NULL;
		if (rc)
			goto out;
	}
	/* a candidate replaces the default only when it is clearly faster:
	 * run-to-run noise of the short timed runs is about 1% */
	if (best_plain < DBL_MAX &&
	    best_plain > r.us_plain_default * (1.0 - HEM_TUNE_MARGIN)) {
		best.ptile_tx = HEM_PTILE_TX_TARGET;
		best.ptile_cls = HEM_PTILE_CLS_TARGET;
		best.ptile_nnz = HEM_PTILE_NNZ_TARGET;
		best_plain = r.us_plain_default;
	}
	r.us_plain_best = best_plain == DBL_MAX ? 0 : best_plain;

	/* batched tiles and chunks (the default tiles were timed above) */
	if (R && best_batch < DBL_MAX) {
		for (size_t i = 1; i < NCAND(batch_cand); ++i) {
			hem_tile_config_default(&g);
			g.tile_tx = batch_cand[i][0];
			g.tile_cls = batch_cand[i][1];
			g.tile_nnz = batch_cand[i][2];
			int imp = 0;
			if ((rc = hem_problem_create_config(cls, &g, &p)))
				goto out;
			if (p->dt.n_tiles)
				rc = time_chunks(p, &t, R, &r, &best_batch,
				                 &best.chunk, &imp, NULL);
			if (!rc && imp) {
				best.tile_tx = g.tile_tx;
				best.tile_cls = g.tile_cls;
				best.tile_nnz = g.tile_nnz;
			}
			hem_problem_destroy(p);
			p = NULL;
			if (rc)
				goto out;
		}
		if (best_batch > r.us_batch_default * (1.0 - HEM_TUNE_MARGIN)) {
			best.tile_tx = HEM_TILE_TX_TARGET;
			best.tile_cls = HEM_TILE_CLS_TARGET;
			best.tile_nnz = HEM_TILE_NNZ_TARGET;
			best.chunk = 0;
			best_batch = r.us_batch_default;
		}
		r.us_batch_best = best_batch;
	}
	*cfg = best;
	if (o->cache_path)
		rc = hem_tune_cache_store(o->cache_path, r.key, &best);
	if (rep)
		*rep = r;
out:
	if (p)
		hem_problem_destroy(p);
	free(t.o1);
	free(t.o2);
	free(t.o3);
	free(t.it);
	free(t.urep);
	return rc;
}
