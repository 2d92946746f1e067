/*
 * problem.c -- validate a class table, transpose it, upload it.
 *
 * The reference walks CLASS->cls[i].ref[k] for every class on every map
 * (src/EM.c:276-286).  Here the table is flattened once into a CSR over
 * classes (members in the caller's order, duplicates kept) and its CSC
 * transpose over transcripts (class ids ascending, a class listed once per
 * occurrence of the transcript), both u32 and device-resident for the life
 * of the handle.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hem_internal.h"

static int validate(const hem_classes *c)
{
	uint32_t i;
	uint64_t nnz;
	if (!c || !c->tx_len || (c->n_cls && (!c->cls_off || !c->cls_count)))
		return hem_fail(HEM_ERR_INVALID, "null class table pointer");
	if (c->n_tx == 0)
		return hem_fail(HEM_ERR_INVALID, "n_tx must be > 0");
	if (c->n_cls && c->cls_off[0] != 0)
		return hem_fail(HEM_ERR_INVALID, "cls_off[0] must be 0");
	for (i = 0; i < c->n_cls; ++i)
		if (c->cls_off[i + 1] < c->cls_off[i])
			return hem_fail(HEM_ERR_INVALID,
			                "cls_off not monotone at class %u", i);
	nnz = c->n_cls ? c->cls_off[c->n_cls] : 0;
	if (nnz && !c->cls_ref)
		return hem_fail(HEM_ERR_INVALID, "null cls_ref");
	for (uint64_t k = 0; k < nnz; ++k)
		if (c->cls_ref[k] >= c->n_tx)
			return hem_fail(HEM_ERR_INVALID,
			                "transcript id %u >= n_tx %u at member %llu",
			                c->cls_ref[k], c->n_tx,
			                (unsigned long long)k);
	for (i = 0; i < c->n_cls; ++i)
		if (!(c->cls_count[i] >= 0) || !isfinite(c->cls_count[i]))
			return hem_fail(HEM_ERR_INVALID,
			                "class %u count is negative or not finite",
			                i);
	for (i = 0; i < c->n_tx; ++i)
		if (c->tx_len[i] == 0)
			return hem_fail(HEM_ERR_INVALID,
			                "transcript %u has length 0", i);
	if (c->member_weight)
		for (uint64_t k = 0; k < nnz; ++k)
			if (!(c->member_weight[k] >= 0) ||
			    !isfinite(c->member_weight[k]))
				return hem_fail(HEM_ERR_INVALID,
				                "member %llu weight is negative or "
				                "not finite",
				                (unsigned long long)k);
	return HEM_OK;
}

static int dalloc(hem_problem *p, void **ptr, size_t n)
{
	int e = hemd_malloc(ptr, n);
	if (e) {
		*ptr = NULL;
		return hem_fail(HEM_ERR_NOMEM, "device allocation of %zu bytes: %s",
		                n, hemd_error_string(e));
	}
	p->dev_bytes += n;
	return HEM_OK;
}

/* Sample layout on the device: {tx0, n_tx, tile0, n_tiles} per sample,
 * tile -> sample, transcript -> sample; host copies of the first
 * transcript and class of each sample. */
static int samples_upload(hem_problem *p, const hem_tiles_host *th,
                          const uint32_t *tx0, const uint32_t *cls0)
{
	const uint32_t ns = p->ns, T = p->n_tx;
	uint32_t *si = calloc((size_t)ns * 4, 4);
	uint32_t *ts = malloc((size_t)T * 4);
	int rc = HEM_OK, e = 0;
	p->h_tx0 = malloc(((size_t)ns + 1) * 4);
	p->h_cls0 = malloc(((size_t)ns + 1) * 4);
	if (!si || !ts || !p->h_tx0 || !p->h_cls0) {
		free(si);
		free(ts);
		return hem_fail(HEM_ERR_NOMEM, "host allocation failed");
	}
	memcpy(p->h_tx0, tx0, ((size_t)ns + 1) * 4);
	memcpy(p->h_cls0, cls0, ((size_t)ns + 1) * 4);
	for (uint32_t q = 0; q < ns; ++q) {
		si[4 * q] = tx0[q];
		si[4 * q + 1] = tx0[q + 1] - tx0[q];
		for (uint32_t t = tx0[q]; t < tx0[q + 1]; ++t)
			ts[t] = q;
	}
	for (uint32_t i = th->n_tiles; i-- > 0;) {
		const uint32_t q = th->tsamp[i];
		si[4 * q + 2] = i;
		si[4 * q + 3]++;
	}
	if ((rc = dalloc(p, (void **)&p->d_sinfo, (size_t)ns * 16)) ||
	    (rc = dalloc(p, (void **)&p->d_tsamp, (size_t)th->n_tiles * 4)) ||
	    (rc = dalloc(p, (void **)&p->d_txsamp, (size_t)T * 4))) {
		free(si);
		free(ts);
		return rc;
	}
	e = hemd_h2d(p->d_sinfo, si, (size_t)ns * 16, p->stream);
	if (!e)
		e = hemd_h2d(p->d_tsamp, th->tsamp, (size_t)th->n_tiles * 4,
		             p->stream);
	if (!e)
		e = hemd_h2d(p->d_txsamp, ts, (size_t)T * 4, p->stream);
	if (!e)
		e = hemd_stream_sync(p->stream);
	free(si);
	free(ts);
	return e ? hem_dev_fail(e, "sample layout upload") : HEM_OK;
}

/* Device problem for a class table holding ns samples side by side
 * (tx0/cls0: first transcript / class of each sample, ns + 1 entries each;
 * NULL with ns == 1). */
/* The single-problem tile kernels' own tiling: the same components packed
 * into larger tiles, and room for the class counts in its tile order
 * (gathered at each solve). */
static int plain_tiles(hem_problem *p, const hem_classes *c,
                       const uint32_t *tx_off, const uint32_t *tx0,
                       uint32_t ns)
{
	hem_tiles_host th;
	int rc = hem_tiles_build(c->n_tx, c->n_cls, c->cls_off, c->cls_ref,
	                         c->tx_len, tx_off, tx0, ns, c->member_weight,
	                         p->cfg.ptile_tx, p->cfg.ptile_cls,
	                         p->cfg.ptile_nnz, &th);
	if (rc == HEM_OK)
		rc = hem_tiles_upload(p, &th, &p->dtp, p->d_ptile_mem,
		                      &p->d_cls_orig_p, &p->ptile_rec_bytes);
	/* per sample: its transcripts and its run of tiles (a sample's tiles
	 * are consecutive; one sample: all of them) */
	if (rc == HEM_OK) {
		uint32_t *pi = calloc((size_t)ns * 4, 4);
		if (!pi)
			rc = hem_fail(HEM_ERR_NOMEM, "host allocation failed");
		for (uint32_t q = 0; rc == HEM_OK && q < ns; ++q) {
			pi[4 * q] = ns > 1 ? tx0[q] : 0;
			pi[4 * q + 1] = ns > 1 ? tx0[q + 1] - tx0[q] : c->n_tx;
		}
		if (rc == HEM_OK && ns == 1)
			pi[3] = th.n_tiles;
		for (uint32_t i = th.n_tiles; rc == HEM_OK && ns > 1 && i-- > 0;) {
			const uint32_t q = th.tsamp[i];
			pi[4 * q + 2] = i;
			pi[4 * q + 3]++;
		}
		if (rc == HEM_OK)
			rc = dalloc(p, (void **)&p->d_ptinfo, (size_t)ns * 16);
		if (rc == HEM_OK &&
		    hemd_h2d(p->d_ptinfo, pi, (size_t)ns * 16, p->stream))
			rc = hem_fail(HEM_ERR_DEVICE, "tile layout upload");
		if (rc == HEM_OK && hemd_stream_sync(p->stream))
			rc = hem_fail(HEM_ERR_DEVICE, "tile layout upload");
		free(pi);
	}
	if (rc == HEM_OK)
		rc = dalloc(p, (void **)&p->d_cntq,
		            (size_t)(th.n_cls ? th.n_cls : 1) * 8);
	/* the monotonicity fallback's per-tile likelihood terms, and the
	 * tree norms' per-tile partials */
	if (rc == HEM_OK)
		rc = dalloc(p, (void **)&p->d_fbpart,
		            (size_t)(th.n_tiles ? th.n_tiles : 1) * 16);
	if (rc == HEM_OK)
		rc = dalloc(p, (void **)&p->d_tpart,
		            (size_t)(th.n_tiles ? th.n_tiles : 1) * 32);
	hem_tiles_free_host(&th);
	return rc;
}

void hem_tile_config_default(hem_tile_config *cfg)
{
	if (!cfg)
		return;
	cfg->ptile_tx = HEM_PTILE_TX_TARGET;
	cfg->ptile_cls = HEM_PTILE_CLS_TARGET;
	cfg->ptile_nnz = HEM_PTILE_NNZ_TARGET;
	cfg->tile_tx = HEM_TILE_TX_TARGET;
	cfg->tile_cls = HEM_TILE_CLS_TARGET;
	cfg->tile_nnz = HEM_TILE_NNZ_TARGET;
	cfg->chunk = 0;
}

static int check_config(const hem_tile_config *g)
{
	const uint32_t ch = g->chunk;
	if (!g->ptile_tx || !g->ptile_cls || !g->ptile_nnz || !g->tile_tx ||
	    !g->tile_cls || !g->tile_nnz)
		return hem_fail(HEM_ERR_INVALID, "tile targets must be > 0");
	if (g->ptile_tx > HEM_TILE_TX || g->tile_tx > HEM_TILE_TX)
		return hem_fail(HEM_ERR_INVALID, "tile transcript targets are at "
		                                 "most %u", HEM_TILE_TX);
	if (g->ptile_cls > HEM_TILE_CLS || g->tile_cls > HEM_TILE_CLS)
		return hem_fail(HEM_ERR_INVALID, "tile class targets are at "
		                                 "most %u", HEM_TILE_CLS);
	if (ch > 32 || (ch & (ch - 1)))
		return hem_fail(HEM_ERR_INVALID, "chunk must be 0 or a power of "
		                                 "two up to 32 (%u)", ch);
	return HEM_OK;
}

static int create_impl(const hem_classes *c, const uint32_t *tx0,
                       const uint32_t *cls0, uint32_t ns,
                       const hem_tile_config *cfg, hem_problem **out)
{
	int rc;
	hem_problem *p;
	uint32_t *tx_off = NULL, *tx_ref = NULL, *pos = NULL;
	double *tx_w = NULL;
	float *cnt32 = NULL;
	double *nd = NULL;
	uint64_t nnz;
	uint32_t T, C, i;

	if (!out)
		return hem_fail(HEM_ERR_INVALID, "null output handle");
	*out = NULL;
	if ((rc = validate(c)))
		return rc;
	if (cfg && (rc = check_config(cfg)))
		return rc;
	if ((rc = hem_device_ready()))
		return rc;
	T = c->n_tx;
	C = c->n_cls;
	nnz = C ? c->cls_off[C] : 0;

	/* CSC transpose by counting sort: class ids come out ascending per
	 * transcript, duplicates adjacent. */
	tx_off = calloc((size_t)T + 1, sizeof(uint32_t));
	pos = malloc((size_t)T * sizeof(uint32_t));
	tx_ref = malloc((nnz ? nnz : 1) * sizeof(uint32_t));
	cnt32 = malloc((C ? C : 1) * sizeof(float));
	nd = calloc((size_t)(C ? C : 1) * 2, sizeof(double));
	if (c->member_weight)
		tx_w = malloc((nnz ? nnz : 1) * sizeof(double));
	if (!tx_off || !pos || !tx_ref || !cnt32 || !nd ||
	    (c->member_weight && !tx_w)) {
		free(tx_off);
		free(pos);
		free(tx_ref);
		free(tx_w);
		free(cnt32);
		free(nd);
		return hem_fail(HEM_ERR_NOMEM, "host allocation failed");
	}
	for (uint64_t k = 0; k < nnz; ++k)
		tx_off[c->cls_ref[k] + 1]++;
	for (i = 0; i < T; ++i)
		tx_off[i + 1] += tx_off[i];
	memcpy(pos, tx_off, (size_t)T * sizeof(uint32_t));
	for (i = 0; i < C; ++i)
		for (uint32_t k = c->cls_off[i]; k < c->cls_off[i + 1]; ++k) {
			if (tx_w)
				tx_w[pos[c->cls_ref[k]]] = c->member_weight[k];
			tx_ref[pos[c->cls_ref[k]]++] = i;
		}
	for (i = 0; i < C; ++i) {
		cnt32[i] = (float)c->cls_count[i];
		nd[2 * (size_t)i] = c->cls_count[i];
	}
	free(pos);

	p = calloc(1, sizeof(*p));
	if (!p) {
		free(tx_off);
		free(tx_ref);
		free(tx_w);
		free(cnt32);
		free(nd);
		return hem_fail(HEM_ERR_NOMEM, "host allocation failed");
	}
	pthread_mutex_init(&p->mu, NULL);
	p->n_tx = T;
	p->n_cls = C;
	p->nnz = nnz;
	p->ns = ns;
	if (cfg)
		p->cfg = *cfg;
	else
		hem_tile_config_default(&p->cfg);
	p->group_cls = 4;
	p->group_tx = 16;

	rc = hemd_stream_create(&p->stream);
	if (!rc)
		rc = hemd_event_create(&p->ev0);
	if (!rc)
		rc = hemd_event_create(&p->ev1);
	if (!rc)
		rc = hemd_event_create(&p->ev2);
	if (!rc)
		rc = hemd_event_create(&p->ev3);
	if (rc) {
		free(tx_off);
		free(tx_ref);
		free(tx_w);
		free(cnt32);
		free(nd);
		hem_problem_destroy(p);
		return hem_dev_fail(rc, "stream/event creation");
	}

#define DALLOC_H(ptr, n)                                           \
	do {                                                       \
		int r_ = dalloc(p, (void **)&(ptr), (n));          \
		if (r_) {                                          \
			free(tx_off);                              \
			free(tx_ref);                              \
			free(tx_w);                                \
			free(cnt32);                               \
			free(nd);                                  \
			hem_problem_destroy(p);                    \
			return r_;                                 \
		}                                                  \
	} while (0)
	DALLOC_H(p->d_cls_off, ((size_t)C + 1) * 4);
	DALLOC_H(p->d_cls_ref, (nnz ? nnz : 1) * 4);
	DALLOC_H(p->d_cnt32, (C ? C : 1) * 4);
	DALLOC_H(p->d_cnt64, (C ? C : 1) * 8);
	DALLOC_H(p->d_nd, (C ? C : 1) * 16);
	DALLOC_H(p->d_tx_off, ((size_t)T + 1) * 4);
	DALLOC_H(p->d_tx_ref, (nnz ? nnz : 1) * 4);
	DALLOC_H(p->d_len, (size_t)T * 4);
	if (tx_w) {
		DALLOC_H(p->d_cls_w, (nnz ? nnz : 1) * 8);
		DALLOC_H(p->d_tx_w, (nnz ? nnz : 1) * 8);
	}
	DALLOC_H(p->d_A, (size_t)T * 8);
	DALLOC_H(p->d_B, (size_t)T * 8);
	DALLOC_H(p->d_C, (size_t)T * 8);
	DALLOC_H(p->d_D, (size_t)T * 8);
	DALLOC_H(p->d_A32, (size_t)T * 4);
	DALLOC_H(p->d_B32, (size_t)T * 4);
	DALLOC_H(p->d_D32, (size_t)T * 4);
	DALLOC_H(p->d_rho, (C ? C : 1) * 8);
	DALLOC_H(p->d_unique, (size_t)T * 4);
	DALLOC_H(p->d_uq, (size_t)T * 4);
	DALLOC_H(p->d_part1, HEMK_NORM_BLOCKS * 2 * 8);
	DALLOC_H(p->d_part2, HEMK_NORM_BLOCKS * 2 * 8);
	DALLOC_H(p->d_out, 64);
	p->spec_L = hemk_spec_chunk(T);
	p->spec_P = (T + p->spec_L - 1) / p->spec_L;
	DALLOC_H(p->d_sx, 2 * (size_t)T * 8);
	DALLOC_H(p->d_scsum, 2 * (size_t)p->spec_P * 8);
	DALLOC_H(p->d_sdrift, 2 * (size_t)p->spec_P * 8);
	DALLOC_H(p->d_scsq, 2 * (size_t)p->spec_P * 8);
	DALLOC_H(p->d_sbase, 2 * (size_t)p->spec_P * 8);
	DALLOC_H(p->d_stab, 2 * (size_t)p->spec_P * HEMK_SPEC_W * 8);
	DALLOC_H(p->d_snx, 2 * (size_t)p->spec_P * HEMK_SPEC_W);
	{
		const size_t nrec = hemk_ws_records(p->spec_L, 2, p->spec_P);
		DALLOC_H(p->d_swa, nrec * 8);
		DALLOC_H(p->d_swlh, nrec * 16);
		DALLOC_H(p->d_swrg, nrec * 4);
		DALLOC_H(p->d_swcnt, 2 * (size_t)p->spec_P * 4);
		DALLOC_H(p->d_scpub, 2 * (size_t)p->spec_P * 8);
	}
	DALLOC_H(p->d_sres, 4 * 8);
	DALLOC_H(p->d_ctl, sizeof(hemk_ctl));
#undef DALLOC_H
	rc = hemd_host_alloc((void **)&p->h_ctl, sizeof(hemk_ctl));
	if (!rc)
		rc = hemd_host_alloc((void **)&p->h_done, 2 * sizeof(uint32_t));
	if (rc) {
		free(tx_off);
		free(tx_ref);
		free(tx_w);
		free(cnt32);
		free(nd);
		hem_problem_destroy(p);
		return hem_dev_fail(rc, "pinned allocation");
	}

	if (C) {
		rc = hemd_h2d(p->d_cls_off, c->cls_off, ((size_t)C + 1) * 4,
		              p->stream);
		if (!rc)
			rc = hemd_h2d(p->d_cls_ref, c->cls_ref, nnz * 4, p->stream);
		if (!rc)
			rc = hemd_h2d(p->d_cnt64, c->cls_count, (size_t)C * 8,
			              p->stream);
		if (!rc)
			rc = hemd_h2d(p->d_cnt32, cnt32, (size_t)C * 4, p->stream);
		if (!rc)
			rc = hemd_h2d(p->d_nd, nd, (size_t)C * 16, p->stream);
		if (!rc)
			rc = hemd_h2d(p->d_tx_ref, tx_ref, nnz * 4, p->stream);
		if (!rc && tx_w && nnz)
			rc = hemd_h2d(p->d_cls_w, c->member_weight, nnz * 8,
			              p->stream);
		if (!rc && tx_w && nnz)
			rc = hemd_h2d(p->d_tx_w, tx_w, nnz * 8, p->stream);
	}
	if (!rc)
		rc = hemd_h2d(p->d_tx_off, tx_off, ((size_t)T + 1) * 4,
		              p->stream);
	if (!rc)
		rc = hemd_h2d(p->d_len, c->tx_len, (size_t)T * 4, p->stream);
	/* the chunk programs' epochs start at 0 */
	if (!rc)
		rc = hemk_fill_u32((uint32_t *)p->d_scpub,
		                   4 * (size_t)p->spec_P, 0, p->stream);

	if (!rc)
		rc = hemd_stream_sync(p->stream);
	if (rc) {
		free(tx_off);
		free(tx_ref);
		free(tx_w);
		free(cnt32);
		free(nd);
		hem_problem_destroy(p);
		return hem_dev_fail(rc, "class table upload");
	}
	/* host copy of the counts (batched solves broadcast them) */
	p->h_cls_count = malloc((C ? C : 1) * sizeof(double));
	if (!p->h_cls_count) {
		free(tx_off);
		free(tx_ref);
		free(tx_w);
		free(cnt32);
		free(nd);
		hem_problem_destroy(p);
		return hem_fail(HEM_ERR_NOMEM, "host allocation failed");
	}
	if (C)
		memcpy(p->h_cls_count, c->cls_count, (size_t)C * sizeof(double));
	p->counts_integral = 1;
	for (i = 0; i < C; ++i)
		if (c->cls_count[i] > 4294967295.0 ||
		    c->cls_count[i] != floor(c->cls_count[i]))
			p->counts_integral = 0;
	/* component tiles for the tile kernels; a table with a component
	 * over the tile limits keeps only the two-pass kernels, and one with
	 * member weights has only the single-problem tiling (weights in its
	 * records) */
	if (c->member_weight) {
		int trc = plain_tiles(p, c, tx_off, NULL, 1);
		if (trc == HEM_OK)
			trc = hemd_stream_sync(p->stream) ? HEM_ERR_DEVICE : HEM_OK;
		if (trc && trc != HEM_ERR_RANGE) {
			free(tx_off);
			free(tx_ref);
			free(tx_w);
			free(cnt32);
			free(nd);
			hem_problem_destroy(p);
			return trc;
		}
	} else {
		hem_tiles_host th;
		int trc = hem_tiles_build(T, C, c->cls_off, c->cls_ref,
		                          c->tx_len, tx_off, tx0, ns, NULL,
		                          p->cfg.tile_tx, p->cfg.tile_cls,
		                          p->cfg.tile_nnz, &th);
		if (trc == HEM_OK) {
			trc = hem_tiles_upload(p, &th, &p->dt, p->d_tile_mem,
			                       &p->d_cls_orig, NULL);
			if (trc == HEM_OK && ns > 1)
				trc = samples_upload(p, &th, tx0, cls0);
			if (trc == HEM_OK)
				trc = plain_tiles(p, c, tx_off, tx0, ns);
			p->h_cls_orig = th.cls_orig;
			th.cls_orig = NULL;
			hem_tiles_free_host(&th);
			if (trc == HEM_OK)
				trc = hemd_stream_sync(p->stream) ? HEM_ERR_DEVICE
				                                  : HEM_OK;
			if (trc) {
				free(tx_off);
				free(tx_ref);
				free(tx_w);
				free(cnt32);
				free(nd);
				hem_problem_destroy(p);
				return trc;
			}
		} else if (trc != HEM_ERR_RANGE || ns > 1) {
			free(tx_off);
			free(tx_ref);
			free(tx_w);
			free(cnt32);
			free(nd);
			hem_problem_destroy(p);
			return trc;
		}
	}
	free(tx_off);
	free(tx_ref);
	free(tx_w);
	free(cnt32);
	free(nd);

	p->dp.n_tx = T;
	p->dp.n_cls = C;
	p->dp.nnz = nnz;
	p->dp.cls_off = p->d_cls_off;
	p->dp.cls_ref = p->d_cls_ref;
	p->dp.cnt32 = p->d_cnt32;
	p->dp.cnt64 = p->d_cnt64;
	p->dp.tx_off = p->d_tx_off;
	p->dp.tx_ref = p->d_tx_ref;
	p->dp.len = p->d_len;
	p->dp.cls_w = p->d_cls_w;
	p->dp.tx_w = p->d_tx_w;
	p->weighted = c->member_weight != NULL;
	*out = p;
	return HEM_OK;
}

int hem_cntq_update(hem_problem *p)
{
	int e;
	if (p->cntq_ok)
		return 0;
	e = hemk_gather_f64(p->d_cnt64, p->d_cls_orig_p, p->dtp.n_cls,
	                    p->d_cntq, p->stream);
	if (!e)
		p->cntq_ok = 1;
	return e;
}

int hem_problem_create(const hem_classes *c, hem_problem **out)
{
	return create_impl(c, NULL, NULL, 1, NULL, out);
}

int hem_problem_create_config(const hem_classes *c,
                              const hem_tile_config *cfg, hem_problem **out)
{
	return create_impl(c, NULL, NULL, 1, cfg, out);
}

int hem_problem_get_config(const hem_problem *p, hem_tile_config *cfg)
{
	if (!p || !cfg)
		return hem_fail(HEM_ERR_INVALID, "null argument");
	*cfg = p->cfg;
	return HEM_OK;
}

int hem_samples_create(const hem_classes *samples, uint32_t n_samples,
                       hem_problem **out)
{
	uint32_t *tx0, *cls0, *off = NULL, *ref = NULL, *len = NULL;
	double *cnt = NULL;
	uint64_t T = 0, C = 0, Z = 0;
	hem_classes all;
	int rc;
	if (!out)
		return hem_fail(HEM_ERR_INVALID, "null output handle");
	*out = NULL;
	if (!samples || n_samples == 0)
		return hem_fail(HEM_ERR_INVALID, "no samples");
	for (uint32_t q = 0; q < n_samples; ++q) {
		if ((rc = validate(&samples[q])))
			return rc;
		if (samples[q].member_weight)
			return hem_fail(HEM_ERR_INVALID, "member weights are for "
			                                 "one-sample handles");
		T += samples[q].n_tx;
		C += samples[q].n_cls;
		Z += samples[q].n_cls ? samples[q].cls_off[samples[q].n_cls] : 0;
	}
	if (T > 0xFFFFFFFFull || C > 0xFFFFFFFEull || Z > 0xFFFFFFFFull)
		return hem_fail(HEM_ERR_RANGE, "samples exceed 2^32 transcripts, "
		                               "classes or members together");
	tx0 = malloc(((size_t)n_samples + 1) * 4);
	cls0 = malloc(((size_t)n_samples + 1) * 4);
	off = malloc((C + 1) * 4);
	ref = malloc((Z ? Z : 1) * 4);
	len = malloc(T * 4);
	cnt = malloc((C ? C : 1) * 8);
	if (!tx0 || !cls0 || !off || !ref || !len || !cnt) {
		rc = hem_fail(HEM_ERR_NOMEM, "host allocation failed");
		goto out;
	}
	/* sample q's transcript ids shift by tx0[q], its classes by cls0[q]:
	 * the table is block diagonal */
	{
		uint32_t t = 0, c = 0, z = 0;
		off[0] = 0;
		for (uint32_t q = 0; q < n_samples; ++q) {
			const hem_classes *sq = &samples[q];
			tx0[q] = t;
			cls0[q] = c;
			memcpy(len + t, sq->tx_len, (size_t)sq->n_tx * 4);
			for (uint32_t i = 0; i < sq->n_cls; ++i) {
				for (uint32_t k = sq->cls_off[i]; k < sq->cls_off[i + 1];
				     ++k)
					ref[z++] = sq->cls_ref[k] + t;
				cnt[c] = sq->cls_count[i];
				off[++c] = z;
			}
			t += sq->n_tx;
		}
		tx0[n_samples] = t;
		cls0[n_samples] = c;
	}
	all.n_tx = (uint32_t)T;
	all.n_cls = (uint32_t)C;
	all.cls_off = off;
	all.cls_ref = ref;
	all.cls_count = cnt;
	all.tx_len = len;
	all.member_weight = NULL;
	rc = create_impl(&all, tx0, cls0, n_samples, NULL, out);
out:
	free(tx0);
	free(cls0);
	free(off);
	free(ref);
	free(len);
	free(cnt);
	return rc;
}

int hem_problem_destroy(hem_problem *p)
{
	if (!p)
		return HEM_OK;
	if (p->stream)
		hemd_stream_sync(p->stream);
	for (int g = 0; g < p->n_graph; ++g)
		hemd_graph_destroy(p->graph[g]);
	hem_batch_free(p);
	hem_batch_destroy_retired(p);
	hem_hera_em_free(p);
	hem_grad_free(p);
	free(p->h_cls_orig);
	free(p->h_cls_count);
	hemd_free(p->d_cls_off);
	hemd_free(p->d_cls_ref);
	hemd_free(p->d_cnt32);
	hemd_free(p->d_cnt64);
	hemd_free(p->d_nd);
	hemd_free(p->d_tx_off);
	hemd_free(p->d_tx_ref);
	hemd_free(p->d_len);
	hemd_free(p->d_A);
	hemd_free(p->d_B);
	hemd_free(p->d_C);
	hemd_free(p->d_D);
	hemd_free(p->d_A32);
	hemd_free(p->d_B32);
	hemd_free(p->d_D32);
	hemd_free(p->d_rho);
	hemd_free(p->d_unique);
	hemd_free(p->d_uq);
	hemd_free(p->d_cntq);
	hemd_free(p->d_sinfo);
	hemd_free(p->d_ptinfo);
	hemd_free(p->d_tsamp);
	hemd_free(p->d_txsamp);
	free(p->h_tx0);
	free(p->h_cls0);
	hemd_free(p->d_part1);
	hemd_free(p->d_part2);
	hemd_free(p->d_out);
	hemd_free(p->d_sx);
	hemd_free(p->d_scsum);
	hemd_free(p->d_sdrift);
	hemd_free(p->d_scsq);
	hemd_free(p->d_sbase);
	hemd_free(p->d_stab);
	hemd_free(p->d_snx);
	hemd_free(p->d_swa);
	hemd_free(p->d_swlh);
	hemd_free(p->d_swrg);
	hemd_free(p->d_swcnt);
	hemd_free(p->d_scpub);
	hemd_free(p->d_sres);
	hemd_free(p->d_fbpart);
	hemd_free(p->d_tpart);
	hemd_free(p->d_cls_w);
	hemd_free(p->d_tx_w);
	for (int i = 0; i < HEM_TILE_ARRAYS; ++i)
		hemd_free(p->d_ptile_mem[i]);
	hemd_free(p->d_ctl);
	hemd_host_free(p->h_ctl);
	hemd_host_free(p->h_done);
	if (p->ev0)
		hemd_event_destroy(p->ev0);
	if (p->ev1)
		hemd_event_destroy(p->ev1);
	if (p->ev2)
		hemd_event_destroy(p->ev2);
	if (p->ev3)
		hemd_event_destroy(p->ev3);
	if (p->stream)
		hemd_stream_destroy(p->stream);
	pthread_mutex_destroy(&p->mu);
	free(p);
	return HEM_OK;
}

int hem_problem_info(const hem_problem *p, uint32_t *n_tx, uint32_t *n_cls,
                     uint64_t *nnz, uint64_t *device_bytes)
{
	if (!p)
		return hem_fail(HEM_ERR_INVALID, "null problem");
	if (n_tx)
		*n_tx = p->n_tx;
	if (n_cls)
		*n_cls = p->n_cls;
	if (nnz)
		*nnz = p->nnz;
	if (device_bytes)
		*device_bytes = p->dev_bytes;
	return HEM_OK;
}

int hem_problem_set_counts(hem_problem *p, const double *cnt)
{
	const uint32_t C = p ? p->n_cls : 0;
	float *c32;
	double *nd;
	int rc = 0, integral = 1;
	uint32_t i;
	if (!p || (C && !cnt))
		return hem_fail(HEM_ERR_INVALID, "null problem or counts");
	for (i = 0; i < C; ++i) {
		if (!(cnt[i] >= 0) || !isfinite(cnt[i]))
			return hem_fail(HEM_ERR_INVALID,
			                "class %u count is negative or not finite",
			                i);
		if (cnt[i] > 4294967295.0 || cnt[i] != floor(cnt[i]))
			integral = 0;
	}
	if (!C)
		return HEM_OK;
	c32 = malloc((size_t)C * sizeof(float));
	nd = malloc((size_t)C * 2 * sizeof(double));
	if (!c32 || !nd) {
		free(c32);
		free(nd);
		return hem_fail(HEM_ERR_NOMEM, "host allocation failed");
	}
	for (i = 0; i < C; ++i) {
		c32[i] = (float)cnt[i];
		nd[2 * (size_t)i] = cnt[i];
		nd[2 * (size_t)i + 1] = 0;
	}
	pthread_mutex_lock(&p->mu);
	p->cntq_ok = 0;
	rc = hemd_h2d(p->d_cnt64, cnt, (size_t)C * 8, p->stream);
	if (!rc)
		rc = hemd_h2d(p->d_cnt32, c32, (size_t)C * 4, p->stream);
	if (!rc)
		rc = hemd_h2d(p->d_nd, nd, (size_t)C * 16, p->stream);
	if (!rc)
		rc = hemd_stream_sync(p->stream);
	if (!rc) {
		memcpy(p->h_cls_count, cnt, (size_t)C * sizeof(double));
		p->counts_integral = integral;
	}
	pthread_mutex_unlock(&p->mu);
	free(c32);
	free(nd);
	return rc ? hem_dev_fail(rc, "count upload") : HEM_OK;
}
