/*
 * hera_em.c -- hera-EM's simplified-model SQUAREM with locking and read
 * filtering (hera-EM/em.c:384-452; see include/hera_em.h) on the device.
 *
 * The filtering state (live members, locked masses, removed classes and
 * their noise banks, locks) lives on the device for the whole solve; the
 * host enqueues one loop at a time and reads back the stop test's count.
 */
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "hem_internal.h"

void hem_hera_em_opts_default(hem_hera_em_opts *o)
{
	if (!o)
		return;
	o->noise_tx = 0xFFFFFFFFu;
	o->max_iter = 1000;
	o->locking = 1;
	o->filtering = 1;
	o->lock_threshold = 1e-5;
	o->filter_threshold = 1e-6;
	o->stop_rel_change = 1e-3;
	o->stop_min_value = 1e-7;
}

static int he_alloc(hem_problem *p)
{
	const size_t T = p->n_tx, C = p->n_cls ? p->n_cls : 1;
	const size_t Z = p->nnz ? p->nnz : 1;
	/* byte flags are cleared and set four at a time */
	const size_t T4 = (T + 3) / 4 * 4, C4 = (C + 3) / 4 * 4;
	const size_t Z4 = (Z + 3) / 4 * 4;
	int e = 0;
#define HA(ptr, bytes)                                                   \
	do {                                                             \
		if (!e)                                                  \
			e = hemd_malloc((void **)&(ptr), (bytes));       \
		if (!e)                                                  \
			p->dev_bytes += (bytes);                         \
	} while (0)
	HA(p->he.occ2mem, Z * 4);
	HA(p->he.counts, 16);
	HA(p->he.lock, T4);
	HA(p->he.alive, Z4);
	HA(p->he.dead, C4);
	HA(p->he.P0, C * 8);
	HA(p->he.bank, C * 8);
	HA(p->he.total, C * 8);
	HA(p->he.brow, C * 8);
	HA(p->he.zrow, (C + 1) * 8);
	HA(p->he.urow, T * 8);
	HA(p->he.lrow, T * 8);
	HA(p->he.sums, 64);
	HA(p->he.p1, T * 8);
	HA(p->he.p2, T * 8);
	HA(p->he.p3, T * 8);
	HA(p->he.acc, T * 8);
#undef HA
	if (e)
		return hem_fail(HEM_ERR_NOMEM, "hera-EM workspace: %s",
		                hemd_error_string(e));
	e = hemk_he_occ2mem(p->n_cls, p->d_cls_off, p->d_cls_ref, p->d_tx_off,
	                    p->d_tx_ref, p->he.occ2mem, p->stream);
	if (e)
		return hem_dev_fail(e, "hera-EM member map");
	p->he.ready = 1;
	return HEM_OK;
}

void hem_hera_em_free(hem_problem *p)
{
	void *a[] = {p->he.occ2mem, p->he.counts, p->he.lock, p->he.alive,
	             p->he.dead, p->he.P0, p->he.bank, p->he.total,
	             p->he.brow, p->he.zrow, p->he.urow, p->he.lrow,
	             p->he.sums, p->he.p1, p->he.p2, p->he.p3, p->he.acc};
	for (size_t i = 0; i < sizeof(a) / sizeof(a[0]); ++i)
		hemd_free(a[i]);
	memset(&p->he, 0, sizeof(p->he));
}

int hem_solve_hera_em(hem_problem *p, const double *p0,
                      const hem_hera_em_opts *opts, double *p_out,
                      hem_hera_em_stats *st)
{
	hem_hera_em_opts o;
	hemk_he h;
	uint32_t k = 0, conv = 0, cnt[4];
	float ms = 0;
	int rc = HEM_OK, e = 0;
	void *s;
	if (!p || !p_out)
		return hem_fail(HEM_ERR_INVALID, "null problem or output");
	if (p->ns > 1)
		return hem_fail(HEM_ERR_INVALID, "one-sample handles only");
	if (opts)
		o = *opts;
	else
		hem_hera_em_opts_default(&o);
	if (!(o.filter_threshold >= 0) || !(o.lock_threshold >= 0) ||
	    !(o.stop_rel_change >= 0) || !(o.stop_min_value >= 0))
		return hem_fail(HEM_ERR_INVALID, "thresholds must be >= 0");
	pthread_mutex_lock(&p->mu);
	s = p->stream;
	if (!p->he.ready && (rc = he_alloc(p)))
		goto out;
	memset(&h, 0, sizeof(h));
	h.T = p->n_tx;
	h.C = p->n_cls;
	h.z = o.noise_tx < p->n_tx ? o.noise_tx : p->n_tx;
	h.thr = o.filter_threshold;
	h.off = p->d_cls_off;
	h.ref = p->d_cls_ref;
	h.tx_off = p->d_tx_off;
	h.tx_ref = p->d_tx_ref;
	h.occ2mem = p->he.occ2mem;
	h.cnt = p->d_cnt64;
	h.w = p->d_cls_w;
	h.tx_w = p->d_tx_w;
	h.lock = p->he.lock;
	h.alive = p->he.alive;
	h.dead = p->he.dead;
	h.P0 = p->he.P0;
	h.bank = p->he.bank;
	h.total = p->he.total;
	h.brow = p->he.brow;
	h.zrow = p->he.zrow;
	h.urow = p->he.urow;
	h.lrow = p->he.lrow;
	h.sums = p->he.sums;
	h.counts = p->he.counts;
	/* a fresh solve: nothing locked, filtered or banked; p1 = p0 or the
	 * uniform start of trans_uniform_init (transdis.c:32-40) */
	if (p0)
		e = hemd_h2d(p->he.p1, p0, (size_t)p->n_tx * 8, s);
	else
		e = hemk_he_fill(p->he.p1, p->n_tx, 1.0 / p->n_tx, s);
	if (!e)
		e = hemk_fill_u32((uint32_t *)p->he.lock, (p->n_tx + 3) / 4, 0,
		                  s);
	if (!e)
		e = hemk_fill_u32((uint32_t *)p->he.dead, (p->n_cls + 3) / 4, 0,
		                  s);
	if (!e)
		e = hemk_fill_u32((uint32_t *)p->he.alive,
		                  ((size_t)p->nnz + 3) / 4, 0x01010101u, s);
	if (!e && p->n_cls)
		e = hemk_he_fill(p->he.P0, p->n_cls, 0.0, s);
	if (!e && p->n_cls)
		e = hemk_he_fill(p->he.bank, p->n_cls, 0.0, s);
	if (!e)
		e = hemk_fill_u32(p->he.counts, 4, 0, s);
	if (!e)
		e = hemd_event_record(p->ev0, s);
	while (!e && k < o.max_iter) {
		double *p1 = p->he.p1, *p2 = p->he.p2, *p3 = p->he.p3;
		double *acc = p->he.acc;
		/* squarem_loop (hera-EM/em.c:384-406) */
		e = hemk_he_estep(&h, p1, 0, acc, s);
		if (!e)
			e = hemk_he_mstep(&h, p1, acc, p2, s);
		if (!e)
			e = hemk_he_estep(&h, p2, o.filtering != 0, acc, s);
		if (!e)
			e = hemk_he_mstep(&h, p2, acc, p3, s);
		if (!e)
			e = hemk_he_accelerate(&h, p1, p2, p3, o.locking != 0,
			                       o.lock_threshold, s);
		if (!e)
			e = hemk_he_estep(&h, p3, 0, acc, s);
		if (!e)
			e = hemk_he_mstep(&h, p3, acc, p1, s);
		if (!e)
			e = hemk_fill_u32(p->he.counts + 2, 1, 0, s);
		if (!e)
			e = hemk_he_stop(&h, p3, p1, o.stop_min_value,
			                 o.stop_rel_change, s);
		if (!e)
			e = hemd_d2h(cnt, p->he.counts, 16, s);
		if (!e)
			e = hemd_stream_sync(s);
		if (e)
			break;
		++k;
		if (cnt[2] == 0) {
			conv = 1;
			break;
		}
	}
	if (!e)
		e = hemd_event_record(p->ev1, s);
	if (!e)
		e = hemd_d2h(p_out, p->he.p1, (size_t)p->n_tx * 8, s);
	if (!e)
		e = hemd_d2h(cnt, p->he.counts, 16, s);
	if (!e)
		e = hemd_stream_sync(s);
	if (!e)
		e = hemd_event_elapsed(p->ev0, p->ev1, &ms);
	if (e) {
		rc = hem_dev_fail(e, "hera-EM solve");
		goto out;
	}
	if (st) {
		uint8_t *lk = malloc(p->n_tx ? p->n_tx : 1);
		memset(st, 0, sizeof(*st));
		st->iters = k;
		st->converged = conv;
		st->removed_members = cnt[0];
		st->removed_classes = cnt[1];
		st->ms = ms;
		if (lk && !hemd_d2h(lk, p->he.lock, p->n_tx, s) &&
		    !hemd_stream_sync(s))
			for (uint32_t t = 0; t < p->n_tx; ++t)
				st->n_locked += lk[t] != 0;
		free(lk);
	}
out:
	pthread_mutex_unlock(&p->mu);
	return rc;
}
