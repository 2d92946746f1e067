/*
 * grad.c -- gradient, Hessian-vector products and standard errors of the
 * abundance log-likelihood (see include/hera_em.h for the objective).
 *
 * The reference has no derivatives; these entry points sit beside the
 * drop-in and reuse its device class table: the gradient is one E-step over
 * the CSR/CSC arrays of the EM kernels, a Hessian-vector product one more
 * pair of passes, and the standard errors a batched conjugate-gradient
 * solve over the component tiles of the batched solver, all directions of
 * one CG step in one tile launch.
 */
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

void hem_grad_free(hem_problem *p)
{
	hem_grad_ws *g = &p->gw;
	void *ptrs[] = {g->theta, g->ell, g->pi,  g->w,   g->x,    g->E,
	                g->fd,    g->gth, g->gell, g->a,  g->v,    g->hv,
	                g->x32,   g->u,   g->rho, g->D,   g->kappa, g->kd,
	                g->kappa_p, g->tot,
	                g->part,  g->scal, g->piq, g->xq, g->aq,   g->mq,
	                g->fdq,   g->kq,  g->jq,  g->pj,  g->X,    g->R,
	                g->P,     g->Q,   g->prz, g->prr, g->ppv,  g->ppfp,
	                g->rz0,   g->rz1, g->rr,  g->bb,  g->pv,   g->pfp,
	                g->pix,   g->se,  g->resid};
	for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i)
		hemd_free(ptrs[i]);
	memset(g, 0, sizeof(*g));
}

#define GALLOC(ptr, bytes)                                                  \
	do {                                                                \
		int e_ = hemd_malloc((void **)&(ptr), (bytes));             \
		if (e_)                                                     \
			return hem_fail(HEM_ERR_NOMEM,                      \
			                "gradient workspace (%zu bytes): %s", \
			                (size_t)(bytes),                    \
			                hemd_error_string(e_));             \
		p->dev_bytes += (bytes);                                    \
	} while (0)

static int ensure_grad(hem_problem *p)
{
	hem_grad_ws *g = &p->gw;
	const size_t T = p->n_tx, C = p->n_cls ? p->n_cls : 1;
	if (g->ready)
		return HEM_OK;
	GALLOC(g->theta, T * 8);
	GALLOC(g->ell, T * 8);
	GALLOC(g->pi, T * 8);
	GALLOC(g->w, T * 8);
	GALLOC(g->x, T * 8);
	GALLOC(g->E, T * 8);
	GALLOC(g->fd, T * 8);
	GALLOC(g->gth, T * 8);
	GALLOC(g->gell, T * 8);
	GALLOC(g->a, T * 8);
	GALLOC(g->v, T * 8);
	GALLOC(g->hv, T * 8);
	GALLOC(g->x32, T * 4);
	GALLOC(g->u, T * 4);
	GALLOC(g->rho, C * 8);
	GALLOC(g->D, C * 8);
	GALLOC(g->kappa, C * 8);
	GALLOC(g->kd, C * 8);
	GALLOC(g->kappa_p, (size_t)(p->dtp.n_cls ? p->dtp.n_cls : 1) * 8);
	{
		/* block partials, or per-tile partials of the tile passes */
		size_t np = HEMK_NORM_BLOCKS;
		if (p->dtp.n_tiles > np)
			np = p->dtp.n_tiles;
		GALLOC(g->part, (3 * np + 3) * 8);
	}
	GALLOC(g->scal, 2 * 8);
	GALLOC(g->tot, (size_t)3 * p->ns * 8);
	g->ready = 1;
	return HEM_OK;
}

static int ensure_se(hem_problem *p, uint32_t K)
{
	hem_grad_ws *g = &p->gw;
	const size_t T = p->n_tx, NC = p->dt.n_cls ? p->dt.n_cls : 1;
	const size_t NRB = hemk_se_rowblocks(p->n_tx);
	if (g->K >= K)
		return HEM_OK;
	{
		void *old[] = {g->piq, g->xq, g->aq, g->mq,  g->fdq, g->kq,
		               g->jq,  g->pj, g->X,  g->R,   g->P,   g->Q,
		               g->prz, g->prr, g->ppv, g->ppfp, g->rz0, g->rz1,
		               g->rr,  g->bb, g->pv, g->pfp, g->pix, g->se,
		               g->resid};
		for (size_t i = 0; i < sizeof(old) / sizeof(old[0]); ++i)
			hemd_free(old[i]);
		g->K = 0;
	}
	GALLOC(g->piq, T * 8);
	GALLOC(g->xq, T * 8);
	GALLOC(g->aq, T * 8);
	GALLOC(g->mq, T * 8);
	GALLOC(g->fdq, T * 8);
	GALLOC(g->kq, NC * 8);
	GALLOC(g->jq, (size_t)K * 4);
	GALLOC(g->pj, (size_t)K * 8);
	GALLOC(g->X, T * K * 8);
	GALLOC(g->R, T * K * 8);
	GALLOC(g->P, T * K * 8);
	GALLOC(g->Q, T * K * 8);
	GALLOC(g->prz, NRB * K * 8);
	GALLOC(g->prr, NRB * K * 8);
	GALLOC(g->ppv, NRB * K * 8);
	GALLOC(g->ppfp, (size_t)(p->dt.n_tiles ? p->dt.n_tiles : 1) * K * 8);
	GALLOC(g->rz0, (size_t)K * 8);
	GALLOC(g->rz1, (size_t)K * 8);
	GALLOC(g->rr, (size_t)K * 8);
	GALLOC(g->bb, (size_t)K * 8);
	GALLOC(g->pv, (size_t)K * 8);
	GALLOC(g->pfp, (size_t)K * 8);
	GALLOC(g->pix, (size_t)K * 8);
	GALLOC(g->se, (size_t)K * 8);
	GALLOC(g->resid, (size_t)K * 8);
	g->K = K;
	return HEM_OK;
}

/* The adjoint and the Hessian-vector product run over the single-problem
 * tiles (classes in LDS) when every component fits one and the tile fits
 * the LDS bound, else over the global CSR/CSC passes. */
static int grad_tiles(const hem_problem *p)
{
	return p->dtp.n_tiles && !p->weighted &&
	       hemk_ptile_lds(&p->dtp) <= HEM_PTILE_LDS_MAX;
}

/* Uploads the point and runs the softmax and the two adjoint passes. */
static int grad_setup(hem_problem *p, const uint32_t *unique,
                      const double *theta, const double *ell, int fp32)
{
	hem_grad_ws *g = &p->gw;
	void *s = p->stream;
	const size_t T = p->n_tx;
	HEM_TRY_DEV(hemd_h2d(g->theta, theta, T * 8, s), "theta upload");
	HEM_TRY_DEV(hemd_h2d(g->ell, ell, T * 8, s), "log_efflen upload");
	HEM_TRY_DEV(hemd_h2d(g->u, unique, T * 4, s), "count upload");
	HEM_TRY_DEV(hemk_grad_softmax(g->theta, g->ell, p->n_tx, g->pi, g->w,
	                              g->x, g->x32, g->part, s),
	            "softmax");
	g->kappa_tiled = 0;
	if (grad_tiles(p)) {
		HEM_TRY_DEV(hem_cntq_update(p), "class count gather");
		HEM_TRY_DEV(hemd_event_record(p->ev2, s), "event");
		HEM_TRY_DEV(hemk_ptile_grad(&p->dtp, p->d_cntq, fp32, g->x,
		                            g->x32, g->u, g->kappa_p, g->E, g->fd,
		                            g->part, s),
		            "adjoint tile passes");
		HEM_TRY_DEV(hemd_event_record(p->ev3, s), "event");
		g->kappa_tiled = 1;
		HEM_TRY_DEV(hemk_grad_fin(g->E, g->pi, g->u, p->n_tx, g->part,
		                          p->dtp.n_tiles, g->gth, g->gell, g->a,
		                          g->fd, g->scal, s),
		            "adjoint end");
		return HEM_OK;
	}
	HEM_TRY_DEV(hemd_event_record(p->ev2, s), "event");
	HEM_TRY_DEV(hemk_grad_passes(&p->dp, fp32, g->x, g->x32, g->u, g->pi,
	                             g->rho, g->D, g->kappa, g->E, g->fd, g->gth,
	                             g->gell, g->a, g->part, g->scal, s),
	            "adjoint passes");
	/* (the global passes' end, k_gfin, falls inside the bracket) */
	HEM_TRY_DEV(hemd_event_record(p->ev3, s), "event");
	return HEM_OK;
}

static int check_point(hem_problem *p, const uint32_t *unique,
                       const double *theta, const double *ell);

static int check_args(hem_problem *p, const uint32_t *unique,
                      const double *theta, const double *ell)
{
	if (p && p->ns > 1)
		return hem_fail(HEM_ERR_INVALID, "one-sample handles only "
		                                 "(hem_grad_samples takes "
		                                 "several)");
	return check_point(p, unique, theta, ell);
}

static int check_point(hem_problem *p, const uint32_t *unique,
                       const double *theta, const double *ell)
{
	if (!p || !unique || !theta || !ell)
		return hem_fail(HEM_ERR_INVALID, "null argument");
	if (p->weighted)
		return hem_fail(HEM_ERR_INVALID, "no likelihood derivatives for "
		                                 "a table with member weights");
	for (uint32_t t = 0; t < p->n_tx; ++t)
		if (!(theta[t] == theta[t]) || theta[t] == 1.0 / 0.0 ||
		    !(ell[t] == ell[t]) || ell[t] == 1.0 / 0.0 ||
		    ell[t] == -1.0 / 0.0)
			return hem_fail(HEM_ERR_INVALID,
			                "theta and log_efflen must be finite "
			                "(theta may be -inf); transcript %u", t);
	return HEM_OK;
}

int hem_grad(hem_problem *p, const uint32_t *unique, const double *theta,
             const double *log_efflen, int precision, double *g_theta,
             double *g_log_efflen, hem_grad_stats *st)
{
	int rc;
	double sc[2];
	float ms = 0;
	if ((rc = check_args(p, unique, theta, log_efflen)))
		return rc;
	if (precision != HEM_PREC_FP32 && precision != HEM_PREC_FP64)
		return hem_fail(HEM_ERR_INVALID, "unknown precision %d",
		                precision);
	pthread_mutex_lock(&p->mu);
	if ((rc = ensure_grad(p)))
		goto out;
	rc = hemd_event_record(p->ev0, p->stream);
	if (rc) {
		rc = hem_dev_fail(rc, "event");
		goto out;
	}
	if ((rc = grad_setup(p, unique, theta, log_efflen,
	                     precision == HEM_PREC_FP32)))
		goto out;
	{
		int e = hemd_event_record(p->ev1, p->stream);
		if (!e && g_theta)
			e = hemd_d2h(g_theta, p->gw.gth, (size_t)p->n_tx * 8,
			             p->stream);
		if (!e && g_log_efflen)
			e = hemd_d2h(g_log_efflen, p->gw.gell,
			             (size_t)p->n_tx * 8, p->stream);
		if (!e)
			e = hemd_d2h(sc, p->gw.scal, 16, p->stream);
		if (!e)
			e = hemd_stream_sync(p->stream);
		if (!e)
			e = hemd_event_elapsed(p->ev0, p->ev1, &ms);
		if (e) {
			rc = hem_dev_fail(e, "gradient readback");
			goto out;
		}
	}
	if (st) {
		float ma = 0;
		if (hemd_event_elapsed(p->ev2, p->ev3, &ma))
			ma = 0;
		st->loglik = sc[0];
		st->n_frag = sc[1];
		st->ms = ms;
		st->ms_adjoint = ma;
	}
out:
	pthread_mutex_unlock(&p->mu);
	return rc;
}

int hem_grad_samples(hem_problem *p, const uint32_t *unique,
                     const double *theta, const double *log_efflen,
                     int precision, double *g_theta, double *g_log_efflen,
                     double *loglik, double *n_frag, hem_grad_stats *st)
{
	hem_grad_ws *g;
	void *s;
	double *tot = NULL;
	float ms = 0, ma = 0;
	int rc;
	const size_t T = p ? p->n_tx : 0;
	if ((rc = check_point(p, unique, theta, log_efflen)))
		return rc;
	if (precision != HEM_PREC_FP32 && precision != HEM_PREC_FP64)
		return hem_fail(HEM_ERR_INVALID, "unknown precision %d",
		                precision);
	if (!grad_tiles(p))
		return hem_fail(HEM_ERR_RANGE, "per-sample gradients run on the "
		                               "component tiles; a component "
		                               "exceeds the tile limits");
	tot = malloc((size_t)3 * p->ns * 8);
	if (!tot)
		return hem_fail(HEM_ERR_NOMEM, "host allocation");
	pthread_mutex_lock(&p->mu);
	g = &p->gw;
	s = p->stream;
	if ((rc = ensure_grad(p)))
		goto out;
	{
		int e = hemd_event_record(p->ev0, s);
		if (!e)
			e = hemd_h2d(g->theta, theta, T * 8, s);
		if (!e)
			e = hemd_h2d(g->ell, log_efflen, T * 8, s);
		if (!e)
			e = hemd_h2d(g->u, unique, T * 4, s);
		if (!e)
			e = hemk_grad_softmax_seg(g->theta, g->ell, p->d_ptinfo,
			                          p->ns, g->pi, g->w, g->x, g->x32,
			                          s);
		if (!e)
			e = hem_cntq_update(p);
		if (!e)
			e = hemd_event_record(p->ev2, s);
		if (!e)
			e = hemk_ptile_grad(&p->dtp, p->d_cntq,
			                    precision == HEM_PREC_FP32, g->x, g->x32,
			                    g->u, g->kappa_p, g->E, g->fd, g->part,
			                    s);
		if (!e)
			e = hemd_event_record(p->ev3, s);
		if (!e)
			e = hemk_grad_fin_seg(g->E, g->pi, g->u, p->n_tx,
			                      p->ns > 1 ? p->d_txsamp : NULL,
			                      p->d_ptinfo, p->ns, g->part,
			                      p->dtp.n_tiles, g->tot, g->gth,
			                      g->gell, g->a, g->fd, s);
		if (!e)
			e = hemd_event_record(p->ev1, s);
		if (!e && g_theta)
			e = hemd_d2h(g_theta, g->gth, T * 8, s);
		if (!e && g_log_efflen)
			e = hemd_d2h(g_log_efflen, g->gell, T * 8, s);
		if (!e)
			e = hemd_d2h(tot, g->tot, (size_t)3 * p->ns * 8, s);
		if (!e)
			e = hemd_stream_sync(s);
		if (!e)
			e = hemd_event_elapsed(p->ev0, p->ev1, &ms);
		if (!e)
			e = hemd_event_elapsed(p->ev2, p->ev3, &ma);
		if (e) {
			rc = hem_dev_fail(e, "per-sample gradient");
			goto out;
		}
	}
	/* the kappa of these passes is per sample; no HVP follows */
	g->kappa_tiled = 1;
	for (uint32_t q = 0; q < p->ns; ++q) {
		if (loglik)
			loglik[q] = tot[3 * q] + tot[3 * q + 2];
		if (n_frag)
			n_frag[q] = tot[3 * q + 1];
	}
	if (st) {
		double L = 0, N = 0;
		for (uint32_t q = 0; q < p->ns; ++q) {
			L += tot[3 * q] + tot[3 * q + 2];
			N += tot[3 * q + 1];
		}
		st->loglik = L;
		st->n_frag = N;
		st->ms = ms;
		st->ms_adjoint = ma;
	}
out:
	pthread_mutex_unlock(&p->mu);
	free(tot);
	return rc;
}

int hem_hvp(hem_problem *p, const uint32_t *unique, const double *theta,
            const double *log_efflen, const double *v, double *hv)
{
	int rc;
	if ((rc = check_args(p, unique, theta, log_efflen)))
		return rc;
	if (!v || !hv)
		return hem_fail(HEM_ERR_INVALID, "null vector");
	pthread_mutex_lock(&p->mu);
	if ((rc = ensure_grad(p)))
		goto out;
	if ((rc = grad_setup(p, unique, theta, log_efflen, 0)))
		goto out;
	{
		hem_grad_ws *g = &p->gw;
		void *s = p->stream;
		int e = hemd_h2d(g->v, v, (size_t)p->n_tx * 8, s);
		if (!e && grad_tiles(p)) {
			e = hemk_dot_parts(g->pi, g->v, p->n_tx, g->part, s);
			if (!e)
				e = hemk_ptile_hvp(&p->dtp, g->x, g->v, g->kappa_p,
				                   g->a, g->pi, g->part, g->scal,
				                   g->hv, s);
		} else if (!e) {
			e = hemk_hvp(&p->dp, g->x, g->a, g->pi, g->kappa, g->scal,
			             g->v, g->kd, g->part, g->hv, s);
		}
		if (!e)
			e = hemd_d2h(hv, g->hv, (size_t)p->n_tx * 8, s);
		if (!e)
			e = hemd_stream_sync(s);
		if (e)
			rc = hem_dev_fail(e, "Hessian-vector product");
	}
out:
	pthread_mutex_unlock(&p->mu);
	return rc;
}

/* pi descending, ties by id */
static const double *g_sort_pi;
static int by_pi(const void *a, const void *b)
{
	const uint32_t i = *(const uint32_t *)a, j = *(const uint32_t *)b;
	if (g_sort_pi[i] != g_sort_pi[j])
		return g_sort_pi[i] > g_sort_pi[j] ? -1 : 1;
	return i < j ? -1 : (i > j);
}
static pthread_mutex_t g_sort_mu = PTHREAD_MUTEX_INITIALIZER;

static int select_top(hem_problem *p, uint32_t K, uint32_t *idx,
                      uint32_t *jq, double *pj)
{
	const uint32_t T = p->n_tx;
	double *pi = malloc((size_t)T * 8);
	uint32_t *ord = malloc((size_t)T * 4), *txq = malloc((size_t)T * 4);
	int e;
	if (!pi || !ord || !txq) {
		free(pi);
		free(ord);
		free(txq);
		return hem_fail(HEM_ERR_NOMEM, "host allocation");
	}
	e = hemd_d2h(pi, p->gw.pi, (size_t)T * 8, p->stream);
	if (!e)
		e = hemd_d2h(txq, p->dt.txq, (size_t)T * 4, p->stream);
	if (!e)
		e = hemd_stream_sync(p->stream);
	if (e) {
		free(pi);
		free(ord);
		free(txq);
		return hem_dev_fail(e, "abundance readback");
	}
	for (uint32_t t = 0; t < T; ++t)
		ord[t] = t;
	pthread_mutex_lock(&g_sort_mu);
	g_sort_pi = pi;
	qsort(ord, T, 4, by_pi);
	pthread_mutex_unlock(&g_sort_mu);
	for (uint32_t k = 0; k < K; ++k) {
		idx[k] = ord[k];
		jq[k] = txq[ord[k]];
		pj[k] = pi[ord[k]];
	}
	free(pi);
	free(ord);
	free(txq);
	return HEM_OK;
}

static int cg_run(hem_problem *p, uint32_t K, uint32_t steps, double damping,
                  int chunk)
{
	hem_grad_ws *g = &p->gw;
	void *s = p->stream;
	const uint32_t T = p->n_tx, NRB = hemk_se_rowblocks(T);
	double *rz = g->rz0, *rzn = g->rz1;
	HEM_TRY_DEV(hemk_cg_init(g->piq, g->mq, g->jq, g->pj, T, K, g->X, g->R,
	                         g->P, g->prz, g->prr, g->ppv, s),
	            "cg init");
	HEM_TRY_DEV(hemk_cg_fold(g->prz, NRB, K, rz, s), "cg fold");
	HEM_TRY_DEV(hemk_cg_fold(g->prr, NRB, K, g->bb, s), "cg fold");
	HEM_TRY_DEV(hemk_cg_fold(g->prr, NRB, K, g->rr, s), "cg fold");
	HEM_TRY_DEV(hemk_cg_fold(g->ppv, NRB, K, g->pv, s), "cg fold");
	for (uint32_t it = 0; it < steps; ++it) {
		HEM_TRY_DEV(hemk_se_tile(&p->dt, chunk, K, g->xq, g->aq, g->kq,
		                         g->piq, g->pv, g->scal, damping, g->P,
		                         g->Q, g->ppfp, s),
		            "cg tile pass");
		HEM_TRY_DEV(hemk_cg_fold(g->ppfp, p->dt.n_tiles, K, g->pfp, s),
		            "cg fold");
		HEM_TRY_DEV(hemk_cg_xr(g->mq, rz, g->pfp, g->rr, g->bb, T, K,
		                       g->P, g->Q, g->X, g->R, g->prz, g->prr, s),
		            "cg update");
		HEM_TRY_DEV(hemk_cg_fold(g->prz, NRB, K, rzn, s), "cg fold");
		HEM_TRY_DEV(hemk_cg_fold(g->prr, NRB, K, g->rr, s), "cg fold");
		HEM_TRY_DEV(hemk_cg_p(g->piq, g->mq, rz, rzn, T, K, g->R, g->P,
		                      g->ppv, s),
		            "cg direction");
		HEM_TRY_DEV(hemk_cg_fold(g->ppv, NRB, K, g->pv, s), "cg fold");
		{
			double *t = rz;
			rz = rzn;
			rzn = t;
		}
	}
	HEM_TRY_DEV(hemk_cg_pix(g->piq, T, K, g->X, g->ppv, s), "cg pi.x");
	HEM_TRY_DEV(hemk_cg_fold(g->ppv, NRB, K, g->pix, s), "cg fold");
	HEM_TRY_DEV(hemk_cg_out(g->X, g->jq, g->pj, g->pix, g->rr, g->bb, K,
	                        g->se, g->resid, s),
	            "cg output");
	return HEM_OK;
}

int hem_stderr(hem_problem *p, const uint32_t *unique, const double *theta,
               const double *log_efflen, uint32_t top_k, uint32_t cg_steps,
               double damping, uint32_t *idx, double *se, double *resid,
               hem_se_stats *st)
{
	int rc, chunk;
	const double t0 = now_ms();
	float ms_setup = 0, ms_cg = 0;
	uint32_t *jq = NULL;
	double *pj = NULL, *rs = NULL;
	if ((rc = check_args(p, unique, theta, log_efflen)))
		return rc;
	if (!idx || !se || top_k == 0 || top_k > p->n_tx)
		return hem_fail(HEM_ERR_INVALID,
		                "top_k must be in 1..n_tx with idx and se given");
	if (!(damping >= 0))
		return hem_fail(HEM_ERR_INVALID, "damping must be >= 0");
	if (!p->dt.n_tiles)
		return hem_fail(HEM_ERR_RANGE,
		                "standard errors need the component tiling");
	chunk = hemk_se_chunk(&p->dt, top_k);
	if (chunk <= 0)
		return hem_fail(HEM_ERR_RANGE, "tiles do not fit the LDS budget");
	jq = malloc((size_t)top_k * 4);
	pj = malloc((size_t)top_k * 8);
	rs = malloc((size_t)top_k * 8);
	if (!jq || !pj || !rs) {
		free(jq);
		free(pj);
		free(rs);
		return hem_fail(HEM_ERR_NOMEM, "host allocation");
	}
	pthread_mutex_lock(&p->mu);
	if ((rc = ensure_grad(p)) || (rc = ensure_se(p, top_k)))
		goto out;
	{
		int e = hemd_event_record(p->ev0, p->stream);
		if (e) {
			rc = hem_dev_fail(e, "event");
			goto out;
		}
	}
	if ((rc = grad_setup(p, unique, theta, log_efflen, 0)))
		goto out;
	if ((rc = select_top(p, top_k, idx, jq, pj)))
		goto out;
	{
		hem_grad_ws *g = &p->gw;
		void *s = p->stream;
		const uint32_t T = p->n_tx;
		int e = hemd_h2d(g->jq, jq, (size_t)top_k * 4, s);
		if (!e)
			e = hemd_h2d(g->pj, pj, (size_t)top_k * 8, s);
		if (!e)
			e = hemk_gather_f64(g->pi, p->dt.tx, T, g->piq, s);
		if (!e)
			e = hemk_gather_f64(g->x, p->dt.tx, T, g->xq, s);
		if (!e)
			e = hemk_gather_f64(g->a, p->dt.tx, T, g->aq, s);
		if (!e)
			e = hemk_gather_f64(g->fd, p->dt.tx, T, g->fdq, s);
		/* kappa by class id from the tile passes' tile order */
		if (!e && g->kappa_tiled)
			e = hemk_scatter_f64(g->kappa_p, p->d_cls_orig_p,
			                     p->dtp.n_cls, g->kappa, s);
		if (!e)
			e = hemk_gather_f64(g->kappa, p->d_cls_orig, p->dt.n_cls,
			                    g->kq, s);
		if (!e)
			e = hemk_precond(g->fdq, g->piq, g->scal, damping, T, g->mq,
			                 s);
		if (!e)
			e = hemd_event_record(p->ev1, s);
		if (e) {
			rc = hem_dev_fail(e, "standard-error setup");
			goto out;
		}
	}
	if ((rc = cg_run(p, top_k, cg_steps, damping, chunk)))
		goto out;
	{
		int e = hemd_event_record(p->ev2, p->stream);
		if (!e)
			e = hemd_d2h(se, p->gw.se, (size_t)top_k * 8, p->stream);
		if (!e)
			e = hemd_d2h(rs, p->gw.resid, (size_t)top_k * 8, p->stream);
		if (!e)
			e = hemd_stream_sync(p->stream);
		if (!e)
			e = hemd_event_elapsed(p->ev0, p->ev1, &ms_setup);
		if (!e)
			e = hemd_event_elapsed(p->ev1, p->ev2, &ms_cg);
		if (e) {
			rc = hem_dev_fail(e, "standard-error readback");
			goto out;
		}
	}
	if (resid)
		memcpy(resid, rs, (size_t)top_k * 8);
	if (st) {
		st->top_k = top_k;
		st->cg_steps = cg_steps;
		st->chunk = (uint32_t)chunk;
		st->resid_max = 0;
		for (uint32_t k = 0; k < top_k; ++k)
			if (rs[k] > st->resid_max)
				st->resid_max = rs[k];
		st->ms_setup = ms_setup;
		st->ms_cg = ms_cg;
		st->ms_total = now_ms() - t0;
	}
out:
	pthread_mutex_unlock(&p->mu);
	free(jq);
	free(pj);
	free(rs);
	return rc;
}
