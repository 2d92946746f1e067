/*
 * em_oracle.c -- CPU restatement of hera's SQUAREM EM (TEST INFRASTRUCTURE).
 *
 * This file is the parity oracle and the timed CPU baseline ("port").  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker; the product library never links it.
 *
 * It restates, in FP64 and with the same loop order and expression order:
 *   M_step          src/EM.c:271-302   (E-step class loop :276-286, M-step
 *                                       and stop test :288-298)
 *   E_step          src/EM.c:304-313   (uniform start 1/T)
 *   estimate_count  src/EM.c:315-387   (SQUAREM loop :325-360, finalise
 *                                       :362-373)
 *   Norm            src/hash_align.c:150-164 (two-pass sample sd)
 * and, behind options that default off:
 *   L2 step length  hera-EM/utility.c:16-41
 *   u2/10 clamp     hera-EM/utility.c:50-53
 *   monotonicity fallback (no counterpart in either reference, which never
 *                   evaluate the likelihood): an iteration whose stabilised
 *                   point lowers the log-likelihood below that of its second
 *                   plain map u2 ends at u2 instead
 * plus three functions the reference lacks (documented in em_oracle.h):
 * the Poisson log-likelihood, its adjoint and its Hessian-vector product.
 *
 * Build: gcc -O2 -fgnu89-inline (the reference's flags, Makefile_linux:9),
 * no fast-math, so the arithmetic is IEEE FP64 without contraction.
 *
 * The compiled reference (src/EM.c) cannot be built here: it includes HDF5
 * headers the image lacks.  The restatement is pinned instead by closed
 * forms, invariants and optimality conditions (tests/test_oracle.py).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "em_oracle.h"

/* The reference's _abs (src/macros.h:13). */
#define O_ABS(x) ((x) < 0 ? -(x) : (x))

typedef struct {
	uint32_t n_tx, n_cls;
	const uint32_t *off, *ref;
	const double *cnt;
	const uint32_t *len;
	const double *w;      /* member weights (CSR order), or NULL */
} o_prob;

/* Norm, src/hash_align.c:150-164: mean, centred sum of squares / (n-1),
 * sqrt.  Sequential sums in index order. */
double oracle_norm_sd(const double *arr, uint32_t n)
{
	uint32_t i;
	double sum, total;

	for (i = 0, sum = 0; i < n; i++)
		sum += arr[i];
	sum /= n;
	for (i = 0, total = 0; i < n; i++)
		total += (arr[i] - sum) * (arr[i] - sum);
	total /= n - 1;
	return sqrt(total);
}

/* One EM map, src/EM.c:271-302.  grama is read, overlap is the scratch the
 * class loop accumulates into and is left zeroed.  Returns a calloc'd
 * vector like the reference (the allocation is part of the timed cost).
 * With member weights the member's share is grama * w in both sums, the
 * responsibility of hera-EM's simple model (s_joint_prob, hera-EM/smodel.h:
 * 15-20, over the read's total, hera-EM/em.c:106-133); w == NULL is the
 * reference's map unchanged. */
static double *o_mstep(const o_prob *P, const double *grama, double *overlap,
                       const uint32_t *count, uint32_t *stop,
                       double stop_min, double stop_rel)
{
	uint32_t i, k;
	double total, max_err, err;
	double *new_grama = calloc(P->n_tx, sizeof(double));

	for (i = 0; i < P->n_cls; ++i) {
		const uint32_t *r = P->ref + P->off[i];
		uint32_t n = P->off[i + 1] - P->off[i];
		const double *w = P->w ? P->w + P->off[i] : NULL;
		if (!w) {
			for (k = 0, total = 0.0; k < n; ++k)
				total += grama[r[k]];
			if (total == 0)
				continue;
			for (k = 0; k < n; ++k)
				overlap[r[k]] += grama[r[k]] * P->cnt[i] / total;
			continue;
		}
		for (k = 0, total = 0.0; k < n; ++k)
			total += grama[r[k]] * w[k];
		if (total == 0)
			continue;
		for (k = 0; k < n; ++k)
			overlap[r[k]] += grama[r[k]] * w[k] * P->cnt[i] / total;
	}

	for (i = 0, max_err = 0.0, *stop = 1; i < P->n_tx; i++) {
		new_grama[i] = (double)(count[i] + overlap[i]) / P->len[i];
		overlap[i] = 0.0;
		if (new_grama[i] > stop_min) {
			err = O_ABS(new_grama[i] - grama[i]) / new_grama[i];
			max_err = err > max_err ? err : max_err;
			if (max_err > stop_rel)
				*stop = 0;
		}
	}
	return new_grama;
}

/* ell(lam) - ell(mu) for the log-likelihood of oracle_loglik (member
 * weights in D_c when the table has them), as sums of per-term differences
 * so that near-equal points do not cancel two large totals: n_c
 * log1p((D_c(lam) - D_c(mu)) / D_c(mu)) over classes (a class with a zero
 * sum at either point contributes its terms as they are, the zero one
 * nothing), u_t log1p((lam_t - mu_t) / mu_t) over transcripts with unique
 * reads, minus len_t (lam_t - mu_t).  Each of the three sums runs in index
 * order; the result is (classes + uniques) - lengths, and *scale receives
 * the same sums of the terms' magnitudes. */
static double o_dloglik(const o_prob *P, const uint32_t *count,
                        const double *lam, const double *mu, double *scale)
{
	uint32_t i, k;
	double a = 0, b = 0, c = 0, sa = 0, sb = 0, sc = 0;
	for (i = 0; i < P->n_cls; ++i) {
		double dl = 0, dm = 0, x = 0;
		if (P->cnt[i] == 0)
			continue;
		for (k = P->off[i]; k < P->off[i + 1]; ++k) {
			const double wk = P->w ? P->w[k] : 1.0;
			dl += lam[P->ref[k]] * wk;
			dm += mu[P->ref[k]] * wk;
		}
		if (dl > 0 && dm > 0)
			x = P->cnt[i] * log1p((dl - dm) / dm);
		else if (dl > 0)
			x = P->cnt[i] * log(dl);
		else if (dm > 0)
			x = -(P->cnt[i] * log(dm));
		a += x;
		sa += fabs(x);
	}
	for (i = 0; i < P->n_tx; ++i) {
		if (count[i] != 0) {
			const double x = count[i] * log1p((lam[i] - mu[i]) / mu[i]);
			b += x;
			sb += fabs(x);
		}
		{
			const double y = P->len[i] * (lam[i] - mu[i]);
			c += y;
			sc += fabs(y);
		}
	}
	if (scale)
		*scale = (sa + sb) + sc;
	return (a + b) - c;
}

/* The fallback's test: the stabilised point lam is kept unless
 * ell(lam) - ell(u2) < -ORACLE_FB_TOL * scale, that is unless it lowers the
 * log-likelihood by more than rounding can explain. */
#define ORACLE_FB_TOL 1e-10

static int o_fallback(const o_prob *P, const uint32_t *count,
                      const double *lam, const double *u2)
{
	double scale = 0;
	const double d = o_dloglik(P, count, lam, u2, &scale);
	return d < -ORACLE_FB_TOL * scale;
}

double oracle_dloglik(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                      const uint32_t *ref, const double *cnt, const double *w,
                      const uint32_t *len, const uint32_t *count,
                      const double *lam, const double *mu, double *scale)
{
	o_prob P = {n_tx, n_cls, off, ref, cnt, len, w};
	return o_dloglik(&P, count, lam, mu, scale);
}

void oracle_opts_default(oracle_opts *o)
{
	o->min_round = 50;
	o->max_round = 1000;
	o->stop_min = 1e-3;
	o->stop_rel = 0.01;
	o->step_form = 0;
	o->clamp_form = 0;
	o->alpha_clamp = 1;
	o->fallback = 0;
}

/* estimate_count, src/EM.c:315-387, on a CSR class table. */
int oracle_estimate_count(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                          const uint32_t *ref, const double *cnt,
                          const uint32_t *len, const uint32_t *count,
                          const oracle_opts *opt, const double *init,
                          double *tpm, double *est, double *lambda,
                          oracle_trace *tr)
{
	return oracle_estimate_count_w(n_tx, n_cls, off, ref, cnt, NULL, len,
	                               count, opt, init, tpm, est, lambda, tr);
}

int oracle_estimate_count_w(uint32_t n_tx, uint32_t n_cls,
                            const uint32_t *off, const uint32_t *ref,
                            const double *cnt, const double *w,
                            const uint32_t *len, const uint32_t *count,
                            const oracle_opts *opt, const double *init,
                            double *tpm, double *est, double *lambda,
                            oracle_trace *tr)
{
	o_prob P = {n_tx, n_cls, off, ref, cnt, len, w};
	uint32_t i, k, stop, stop2, fb;
	double *u, *u1, *u2, *grama, *overlap, sum, alpha, n_v, n_r;
	double *r = malloc(n_tx * sizeof(double));
	double *v = malloc(n_tx * sizeof(double));

	if (!r || !v || n_tx == 0)
		return 1;
	/* E_step, src/EM.c:304-313 */
	grama = calloc(n_tx, sizeof(double));
	overlap = calloc(n_tx, sizeof(double));
	for (i = 0; i < n_tx; i++)
		grama[i] = init ? init[i] : (double)1 / n_tx;

	stop = 0;
	k = 0;
	while (k < opt->min_round || (k < opt->max_round && stop == 0)) {
		u = grama;
		u1 = o_mstep(&P, u, overlap, count, &stop, opt->stop_min,
		             opt->stop_rel);
		u2 = o_mstep(&P, u1, overlap, count, &stop, opt->stop_min,
		             opt->stop_rel);
		stop2 = stop;

		for (i = 0; i < n_tx; ++i) {
			r[i] = u1[i] - u[i];
			v[i] = (u2[i] - u1[i]) - r[i];
		}

		if (opt->step_form == 0) {
			n_r = oracle_norm_sd(r, n_tx);
			n_v = oracle_norm_sd(v, n_tx);
		} else {
			/* hera-EM/utility.c:34-41: squared L2 norms; a zero
			 * ||v|| (<= 1e-300) gives gamma^2 = 1. */
			double r2 = 0, v2 = 0;
			for (i = 0; i < n_tx; ++i) {
				r2 += r[i] * r[i];
				v2 += v[i] * v[i];
			}
			n_r = sqrt(r2);
			n_v = v2 <= 1e-300 ? 0 : sqrt(v2);
		}

		if (n_v != 0)
			alpha = -n_r / n_v;
		else
			alpha = -1;
		if (opt->alpha_clamp && alpha > -1)
			alpha = -1;

		for (i = 0; i < n_tx; ++i) {
			double x = u[i] - 2 * alpha * r[i] + alpha * alpha * v[i];
			if (opt->clamp_form == 0) {
				if (x < 0)
					x = 0;
			} else if (x <= 1e-300) {
				x = u2[i] / 10;
			}
			u1[i] = x;
		}

		grama = o_mstep(&P, u1, overlap, count, &stop, opt->stop_min,
		                opt->stop_rel);
		fb = 0;
		if (opt->fallback && o_fallback(&P, count, grama, u2)) {
			/* the stabilised point is worse than two plain maps:
			 * keep u2, with the stop test of the map that made it */
			memcpy(grama, u2, n_tx * sizeof(double));
			stop = stop2;
			fb = 1;
		}
		free(u2);
		if (tr && tr->fallback && k < tr->cap)
			tr->fallback[k] = fb;
		if (tr && k < tr->cap) {
			tr->alpha[k] = alpha;
			tr->stop[k] = stop;
		}
		if (tr && tr->last_u)
			memcpy(tr->last_u, u1, n_tx * sizeof(double));
		free(u);
		free(u1);
		++k;
	}

	if (lambda)
		memcpy(lambda, grama, n_tx * sizeof(double));
	/* finalisation, src/EM.c:362-373 */
	for (i = 0, sum = 0.0; i < n_tx; ++i)
		sum += grama[i];
	for (i = 0; i < n_tx; ++i) {
		double e = 0, g = 0;
		if (sum != 0) {
			e = grama[i] * len[i];
			g = grama[i] * 1000000 / sum;
		}
		if (est)
			est[i] = e;
		if (tpm)
			tpm[i] = g;
	}
	if (tr)
		tr->k = k;
	free(grama);
	free(overlap);
	free(r);
	free(v);
	return 0;
}

/* One EM map, exposed for closed-form tests of the class loop. */
int oracle_em_map(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                  const uint32_t *ref, const double *cnt, const uint32_t *len,
                  const uint32_t *count, const double *lam_in,
                  double *lam_out, uint32_t *stop)
{
	return oracle_em_map_w(n_tx, n_cls, off, ref, cnt, NULL, len, count,
	                       lam_in, lam_out, stop);
}

int oracle_em_map_w(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                    const uint32_t *ref, const double *cnt, const double *w,
                    const uint32_t *len, const uint32_t *count,
                    const double *lam_in, double *lam_out, uint32_t *stop)
{
	o_prob P = {n_tx, n_cls, off, ref, cnt, len, w};
	double *ov = calloc(n_tx, sizeof(double));
	double *nw;
	if (!ov)
		return 1;
	nw = o_mstep(&P, lam_in, ov, count, stop, 1e-3, 0.01);
	memcpy(lam_out, nw, n_tx * sizeof(double));
	free(nw);
	free(ov);
	return 0;
}

/* Poisson log-likelihood whose EM map is src/EM.c:271-302:
 *   l = sum_t u_t log l_t + sum_{c: D_c > 0} n_c log D_c - sum_t len_t l_t.
 * A class whose members are all exactly zero (the extrapolation clamp of
 * src/EM.c:351 can do this) is skipped by the map (src/EM.c:279-280) and
 * contributes nothing here either. */
double oracle_loglik(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                     const uint32_t *ref, const double *cnt,
                     const uint32_t *len, const uint32_t *count,
                     const double *lam)
{
	uint32_t i, k;
	double a = 0, b = 0, c = 0;
	for (i = 0; i < n_cls; ++i) {
		double d = 0;
		for (k = off[i]; k < off[i + 1]; ++k)
			d += lam[ref[k]];
		if (cnt[i] != 0 && d > 0)
			a += cnt[i] * log(d);
	}
	for (i = 0; i < n_tx; ++i) {
		if (count[i] != 0)
			b += count[i] * log(lam[i]);
		c += len[i] * lam[i];
	}
	return a + b - c;
}

/* ---- adjoint of the softmax-parameterised likelihood --------------------
 * pi = softmax(theta), w_t = exp(-ell_t) (ell = log effective length),
 * unique reads as singleton classes:
 *   L = sum_c n_c log D_c + sum_t u_t log(pi_t w_t),  D_c = sum pi_s w_s.
 * With s_t = sum_{c ni t} n_c / D_c + u_t / (pi_t w_t) and E_t = pi_t w_t s_t
 * (expected fragments of t), N = sum_t E_t:
 *   dL/dtheta_j = E_j - N pi_j,   dL/dell_j = -E_j.                       */
static void o_softmax(uint32_t n, const double *theta, double *pi)
{
	uint32_t i;
	double m = theta[0], s = 0;
	for (i = 1; i < n; ++i)
		if (theta[i] > m)
			m = theta[i];
	for (i = 0; i < n; ++i)
		s += (pi[i] = exp(theta[i] - m));
	for (i = 0; i < n; ++i)
		pi[i] /= s;
}

double oracle_grad(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                   const uint32_t *ref, const double *cnt,
                   const uint32_t *count, const double *theta,
                   const double *ell, double *g_theta, double *g_ell)
{
	uint32_t i, k;
	double L = 0, N = 0;
	double *pi = malloc(n_tx * sizeof(double));
	double *x = malloc(n_tx * sizeof(double));
	double *E = calloc(n_tx, sizeof(double));
	o_softmax(n_tx, theta, pi);
	for (i = 0; i < n_tx; ++i)
		x[i] = pi[i] * exp(-ell[i]);
	for (i = 0; i < n_cls; ++i) {
		double d = 0;
		for (k = off[i]; k < off[i + 1]; ++k)
			d += x[ref[k]];
		if (cnt[i] == 0 || d == 0)
			continue;
		L += cnt[i] * log(d);
		for (k = off[i]; k < off[i + 1]; ++k)
			E[ref[k]] += x[ref[k]] * (cnt[i] / d);
	}
	for (i = 0; i < n_tx; ++i) {
		if (count[i]) {
			L += count[i] * log(x[i]);
			E[i] += count[i];
		}
	}
	for (i = 0; i < n_tx; ++i)
		N += E[i];
	for (i = 0; i < n_tx; ++i) {
		if (g_theta)
			g_theta[i] = E[i] - N * pi[i];
		if (g_ell)
			g_ell[i] = -E[i];
	}
	free(pi);
	free(x);
	free(E);
	return L;
}

/* Hessian-vector product d/de [grad_theta L](theta + e*vec) at e=0:
 *   pid_t = pi_t (vec_t - pi.vec),  Dd_c = sum pid_s w_s,
 *   sd_t  = -sum_{c ni t} n_c Dd_c / D_c^2   (unique reads: singleton),
 *   Ed_t  = pid_t w_t s_t + pi_t w_t sd_t,   Hv_t = Ed_t - N pid_t.        */
void oracle_hvp(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                const uint32_t *ref, const double *cnt,
                const uint32_t *count, const double *theta, const double *ell,
                const double *vec, double *hv)
{
	uint32_t i, k;
	double N = 0, pv = 0;
	double *pi = malloc(n_tx * sizeof(double));
	double *w = malloc(n_tx * sizeof(double));
	double *pd = malloc(n_tx * sizeof(double));
	double *s = calloc(n_tx, sizeof(double));
	double *sd = calloc(n_tx, sizeof(double));
	o_softmax(n_tx, theta, pi);
	for (i = 0; i < n_tx; ++i) {
		w[i] = exp(-ell[i]);
		pv += pi[i] * vec[i];
	}
	for (i = 0; i < n_tx; ++i)
		pd[i] = pi[i] * (vec[i] - pv);
	for (i = 0; i < n_cls; ++i) {
		double d = 0, dd = 0;
		for (k = off[i]; k < off[i + 1]; ++k) {
			d += pi[ref[k]] * w[ref[k]];
			dd += pd[ref[k]] * w[ref[k]];
		}
		if (cnt[i] == 0 || d == 0)
			continue;
		N += cnt[i];
		for (k = off[i]; k < off[i + 1]; ++k) {
			s[ref[k]] += cnt[i] / d;
			sd[ref[k]] -= cnt[i] * dd / (d * d);
		}
	}
	for (i = 0; i < n_tx; ++i) {
		if (count[i]) {
			double d = pi[i] * w[i], dd = pd[i] * w[i];
			N += count[i];
			s[i] += count[i] / d;
			sd[i] -= count[i] * dd / (d * d);
		}
	}
	for (i = 0; i < n_tx; ++i)
		hv[i] = pd[i] * w[i] * s[i] + pi[i] * w[i] * sd[i] - N * pd[i];
	free(pi);
	free(w);
	free(pd);
	free(s);
	free(sd);
}

/* ---- bootstrap baseline: the reference's join-per-batch thread schedule
 * (src/EM.c:243-255), each thread running one estimate_count on its own
 * count vector.  Used only to time the CPU baseline. */
typedef struct {
	const o_prob *P;
	const uint32_t *count;
	const double *cnt;
	const oracle_opts *opt;
	double *est;
	uint32_t k;
} o_job;

static void *o_bs_thread(void *arg)
{
	o_job *j = (o_job *)arg;
	oracle_trace tr;
	memset(&tr, 0, sizeof(tr));
	oracle_estimate_count(j->P->n_tx, j->P->n_cls, j->P->off, j->P->ref,
	                      j->cnt, j->P->len, j->count, j->opt, NULL, NULL,
	                      j->est, NULL, &tr);
	j->k = tr.k;
	return NULL;
}

int oracle_bootstrap_batch(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                           const uint32_t *ref, const uint32_t *len,
                           uint32_t n_rep, const uint32_t *counts,
                           const double *cls_counts, const oracle_opts *opt,
                           uint32_t n_thread, double *est, uint32_t *k_out)
{
	o_prob P = {n_tx, n_cls, off, ref, NULL, len};
	pthread_t *thr = calloc(n_thread, sizeof(pthread_t));
	o_job *jobs = calloc(n_rep, sizeof(o_job));
	pthread_attr_t attr;
	uint32_t i, k;

	if (!thr || !jobs)
		return 1;
	pthread_attr_init(&attr);
	pthread_attr_setstacksize(&attr, 16u << 20);
	for (i = 0; i < n_rep;) {
		for (k = 0; k < n_thread && i < n_rep; ++k, ++i) {
			jobs[i].P = &P;
			jobs[i].count = counts + (size_t)i * n_tx;
			jobs[i].cnt = cls_counts + (size_t)i * n_cls;
			jobs[i].opt = opt;
			jobs[i].est = est ? est + (size_t)i * n_tx : NULL;
			pthread_create(&thr[k], &attr, o_bs_thread, &jobs[i]);
		}
		for (; k > 0; --k)
			pthread_join(thr[k - 1], NULL);
	}
	for (i = 0; i < n_rep; ++i)
		if (k_out)
			k_out[i] = jobs[i].k;
	pthread_attr_destroy(&attr);
	free(thr);
	free(jobs);
	return 0;
}

double oracle_now(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* ---- bootstrap resampling restatement ------------------------------------
 * The library draws replicate counts as Multinomial(N, w/N) over K
 * categories in two levels (hera_amd/csrc/dev/resample.hip): level 1 sends
 * each of the N fragments to a chunk of categories, level 2 spreads each
 * chunk's fragments over its categories.  Restated here sequentially,
 * from the published generator (Philox-4x32-10, Salmon et al., SC'11) and
 * Lemire's multiply-and-reject bounded integers, to check the device
 * counts bit for bit. */
void oracle_philox(uint64_t key, const uint32_t ctr[4], uint32_t out[4])
{
	uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
	uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
	int r;
	for (r = 0; r < 10; ++r) {
		uint64_t a = (uint64_t)0xD2511F53u * c0;
		uint64_t b = (uint64_t)0xCD9E8D57u * c2;
		uint32_t n0 = (uint32_t)(b >> 32) ^ c1 ^ k0;
		uint32_t n2 = (uint32_t)(a >> 32) ^ c3 ^ k1;
		c1 = (uint32_t)b;
		c3 = (uint32_t)a;
		c0 = n0;
		c2 = n2;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	out[0] = c0;
	out[1] = c1;
	out[2] = c2;
	out[3] = c3;
}

static uint32_t o_bounded(uint64_t key, uint32_t i, uint32_t j, uint32_t M)
{
	uint32_t thresh = (0u - M) % M, att, w;
	for (att = 0;; ++att) {
		uint32_t ctr[4] = {i, j, att, 0x62737472u}, h[4];
		oracle_philox(key, ctr, h);
		for (w = 0; w < 4; ++w) {
			uint64_t m = (uint64_t)h[w] * M;
			if ((uint32_t)m >= thresh)
				return (uint32_t)(m >> 32);
		}
	}
}

/* last index j in [0, n) with a[j] <= y, for a[0] <= y < a[n] */
static uint32_t o_last_le(const uint32_t *a, uint32_t n, uint32_t y)
{
	uint32_t lo = 0, hi = n;
	while (hi - lo > 1) {
		uint32_t mid = (lo + hi) / 2;
		if (a[mid] <= y)
			lo = mid;
		else
			hi = mid;
	}
	return lo;
}

#define O_RS_CHUNK 2048
#define O_RS_MAXB 7000

/* Chunk plan: a category heavier than the target weight W is a chunk on its
 * own; other categories fill the open chunk until it would pass W or hold
 * O_RS_CHUNK categories; W = ceil(N/2048), doubled until <= O_RS_MAXB
 * chunks. */
uint32_t oracle_resample_plan(const uint32_t *w, uint32_t K, uint32_t N,
                              uint32_t *start, uint32_t *cw)
{
	uint64_t W = ((uint64_t)N + 2047) / 2048;
	if (W == 0)
		W = 1;
	for (;;) {
		uint32_t B = 0, n = 0, k;
		uint64_t cur = 0, acc = 0;
		for (k = 0; k < K; ++k) {
			if (w[k] > W) {
				if (n) {
					cw[B] = (uint32_t)acc;
					start[B++] = k - n;
					acc += cur;
					n = 0;
					cur = 0;
				}
				cw[B] = (uint32_t)acc;
				start[B++] = k;
				acc += w[k];
			} else {
				if (n && (cur + w[k] > W || n == O_RS_CHUNK)) {
					cw[B] = (uint32_t)acc;
					start[B++] = k - n;
					acc += cur;
					n = 0;
					cur = 0;
				}
				cur += w[k];
				++n;
			}
			if (B > O_RS_MAXB)
				break;
		}
		if (k == K && n) {
			cw[B] = (uint32_t)acc;
			start[B++] = K - n;
			acc += cur;
		}
		if (k == K && B <= O_RS_MAXB) {
			start[B] = K;
			cw[B] = (uint32_t)acc;
			return B;
		}
		W *= 2;
	}
}

/* One replicate's counts out[K] for weights w (N = sum w < 2^32). */
int oracle_resample(const uint32_t *w, uint32_t K, uint64_t seed,
                    uint32_t *out)
{
	uint32_t *start = malloc((O_RS_MAXB + 2) * 4);
	uint32_t *cw = malloc((O_RS_MAXB + 2) * 4);
	uint32_t *m, *cdf, B, i, j, k;
	uint64_t N64 = 0;
	for (k = 0; k < K; ++k)
		N64 += w[k];
	if (!start || !cw || N64 > 0xFFFFFFFFull)
		return 1;
	B = oracle_resample_plan(w, K, (uint32_t)N64, start, cw);
	m = calloc(B ? B : 1, 4);
	cdf = malloc((O_RS_CHUNK + 1) * 4);
	for (i = 0; i < (uint32_t)N64; ++i)
		m[o_last_le(cw, B, o_bounded(seed, i, 0xFFFFFFFFu,
		                             (uint32_t)N64))]++;
	memset(out, 0, (size_t)K * 4);
	for (j = 0; j < B; ++j) {
		uint32_t k0 = start[j], nk = start[j + 1] - start[j], acc = 0;
		if (nk == 1) {
			out[k0] = m[j];
			continue;
		}
		for (k = 0; k < nk; ++k) {
			cdf[k] = acc;
			acc += w[k0 + k];
		}
		cdf[nk] = acc;
		for (i = 0; i < m[j]; ++i)
			out[k0 + o_last_le(cdf, nk, o_bounded(seed, i, j, acc))]++;
	}
	free(start);
	free(cw);
	free(m);
	free(cdf);
	return 0;
}

/* ---- timed CPU bootstrap: hera's schedule with the multinomial resample --
 * Batches of n_thread threads joined per batch (src/EM.c:243-255); each
 * thread resamples its replicate (oracle_resample over unique reads then
 * classes) and runs estimate_count on the new counts. */
typedef struct {
	uint32_t n_tx, n_cls;
	const uint32_t *off, *ref, *len, *w;
	const oracle_opts *opt;
	uint64_t seed;
	double *est;
	uint32_t k;
} o_bjob;

static void *o_boot_thread(void *arg)
{
	o_bjob *j = (o_bjob *)arg;
	uint32_t K = j->n_tx + j->n_cls, c;
	uint32_t *cnt = malloc((size_t)K * 4);
	double *cc = malloc(((size_t)j->n_cls + 1) * sizeof(double));
	oracle_trace tr;
	memset(&tr, 0, sizeof(tr));
	oracle_resample(j->w, K, j->seed, cnt);
	for (c = 0; c < j->n_cls; ++c)
		cc[c] = cnt[j->n_tx + c];
	oracle_estimate_count(j->n_tx, j->n_cls, j->off, j->ref, cc, j->len,
	                      cnt, j->opt, NULL, NULL, j->est, NULL, &tr);
	j->k = tr.k;
	free(cnt);
	free(cc);
	return NULL;
}

int oracle_bootstrap_resampled(uint32_t n_tx, uint32_t n_cls,
                               const uint32_t *off, const uint32_t *ref,
                               const uint32_t *len, const uint32_t *w,
                               uint32_t n_rep, const uint64_t *seeds,
                               const oracle_opts *opt, uint32_t n_thread,
                               double *est, uint32_t *k_out)
{
	pthread_t *thr = calloc(n_thread, sizeof(pthread_t));
	o_bjob *jobs = calloc(n_rep, sizeof(o_bjob));
	pthread_attr_t attr;
	uint32_t i, k;
	if (!thr || !jobs)
		return 1;
	pthread_attr_init(&attr);
	pthread_attr_setstacksize(&attr, 16u << 20);
	for (i = 0; i < n_rep;) {
		for (k = 0; k < n_thread && i < n_rep; ++k, ++i) {
			o_bjob *j = &jobs[i];
			j->n_tx = n_tx;
			j->n_cls = n_cls;
			j->off = off;
			j->ref = ref;
			j->len = len;
			j->w = w;
			j->opt = opt;
			j->seed = seeds[i];
			j->est = est ? est + (size_t)i * n_tx : NULL;
			pthread_create(&thr[k], &attr, o_boot_thread, j);
		}
		for (; k > 0; --k)
			pthread_join(thr[k - 1], NULL);
	}
	if (k_out)
		for (i = 0; i < n_rep; ++i)
			k_out[i] = jobs[i].k;
	pthread_attr_destroy(&attr);
	free(thr);
	free(jobs);
	return 0;
}

/* ---- abundance tables: src/EM.c:176-205 restated ------------------------
 * abundance.tsv: the header of :178, then per transcript the fprintf of
 * :180-184 with the name block as a C string.  abundance.gene.tsv: the
 * header of :188, then the loop of :189-204 -- float sums, the row named by
 * the last transcript of the run (name block + 16) and the gene's name
 * block + 16 -- except that the run also ends at the last transcript
 * (the reference compares gene[n], one past its array, there). */
int oracle_write_abundance(const char *path, uint32_t n, const char *names,
                           uint32_t name_len, const uint32_t *count,
                           const uint32_t *len, const double *eff_len,
                           const double *overlap, const double *grama)
{
	FILE *out = fopen(path, "w");
	uint32_t i;
	if (!out)
		return 1;
	fprintf(out, "#target_id\tunique_map\tlength\teff_length\test_counts\ttpm\n");
	for (i = 0; i < n; ++i)
		fprintf(out, "%s\t%u\t%u\t%f\t%f\t%f\n", names + i * name_len,
		        count[i], len[i], eff_len[i], overlap[i], grama[i]);
	return fclose(out) != 0;
}

int oracle_write_gene_abundance(const char *path, uint32_t n,
                                const char *names, uint32_t name_len,
                                const uint32_t *gene, const char *gene_name,
                                uint32_t l_gene, const double *overlap,
                                const double *grama)
{
	FILE *out = fopen(path, "w");
	uint32_t i, g;
	float sum[2];
	if (!out)
		return 1;
	fprintf(out, "#gene_id\tgene_name\test_counts\ttpm\n");
	for (i = 0; i < n; ++i) {
		g = gene[i];
		sum[0] = overlap[i];
		sum[1] = grama[i];
		while (i + 1 < n && g == gene[i + 1]) {
			++i;
			sum[0] += overlap[i];
			sum[1] += grama[i];
		}
		fprintf(out, "%s\t%s\t%f\t%f\n", names + i * name_len + 16,
		        gene_name + g * l_gene + 16, sum[0], sum[1]);
	}
	return fclose(out) != 0;
}

/* ---- hera's bootstrap sampler: src/EM.c:98-150 restated -----------------
 * oracle_rand_r is glibc's rand_r (three steps of the LCG
 * x -> 1103515245 x + 12345 mod 2^32; 11 bits of bits 16.. of the first
 * state, then 10 and 10 of the next two, shifted in).  oracle_hera_sample
 * builds DIST->table as create_distribution does (one entry per unique
 * read, in transcript order) and runs bootstrap_sample's loop
 * `n = N; while (--n > 0) ++count[table[rand_r(&seed) % n]];` from *seed,
 * leaving *seed where the loop leaves it (hera's global seed). */
uint32_t oracle_rand_r(uint32_t *seed)
{
	uint32_t next = *seed, result;
	next = next * 1103515245u + 12345u;
	result = (next / 65536u) % 2048u;
	next = next * 1103515245u + 12345u;
	result <<= 10;
	result ^= (next / 65536u) % 1024u;
	next = next * 1103515245u + 12345u;
	result <<= 10;
	result ^= (next / 65536u) % 1024u;
	*seed = next;
	return result;
}

int oracle_hera_sample(const uint32_t *unique, uint32_t n_tx, uint32_t *seed,
                       uint32_t *count)
{
	uint64_t N = 0;
	uint32_t i, k, n, *table;
	for (i = 0; i < n_tx; ++i)
		N += unique[i];
	memset(count, 0, (size_t)n_tx * sizeof(uint32_t));
	if (N == 0)
		return 0;
	if (N > 0xFFFFFFFFull)
		return 1;
	table = malloc(N * sizeof(uint32_t));
	if (!table)
		return 1;
	for (i = 0, k = 0; i < n_tx; ++i) {
		uint32_t c;
		for (c = 0; c < unique[i]; ++c)
			table[k++] = i;
	}
	n = (uint32_t)N;
	while (--n > 0)
		++count[table[oracle_rand_r(seed) % n]];
	free(table);
	return 0;
}

/* ---- hera-EM's simplified-model SQUAREM --------------------------------
 * s_squarem (hera-EM/em.c:418-426) over squarem_loop (em.c:384-406) with
 * the simple model's E-step (the e_step macro, em.c:75-157, with
 * s_joint_prob / s_noise_prob / s_model_update / s_model_update_noise,
 * smodel.h:15-45), M-step (trans_master_init + lock_normalize,
 * transdis.c:20-30, 54-61, utility.h:104-122), locked extrapolation
 * (lock_accelerate + the locking rule, utility.c:64-107, transdis.c:
 * 97-115) and stop test (is_stop / arr_diff, transdis.c:123-130,
 * utility.c:109-128), on the class table: a class is a set of identical
 * reads (count n_c) whose alignments are its members (transcript t_k,
 * weight w_k = c_prob); the noise component is transcript noise_tx, a
 * member of every class with weight 4^-read_length.
 *
 * Restated per class, in class order (hera-EM's per-thread joins have no
 * fixed order):
 *   E-step (p -> acc), flag filt (the second EM of each loop):
 *     acc[noise] starts at the banked noise count B (sum over removed
 *     classes of their banks, class order) when B >= thr and the noise is
 *     not locked (n_update(new_m, NULL, noise_cnt), em.c:85);
 *     per live class: total = P0_c + sum over live members of p_t w_k
 *     (member order); below 1e-300 the class is skipped (and, filtering,
 *     removed with bank n_c, em.c:112-115);
 *     per live member: post = (p_t w_k) / total; the noise member adds
 *     n_c post to acc[noise] unless post < thr or the noise is locked;
 *     another member adds n_c post to acc[t] unless post < thr (filtering:
 *     the member is removed) or t is locked (filtering: removed, and
 *     P0_c += p_t w_k, member order);
 *     filtering, a class left without a contributing member is removed
 *     and banks n_c times its noise posterior (em.c:139-145).
 *   M-step: locked entries keep p; the others are acc scaled by
 *     (1 / sum of unlocked acc) * (1 - sum of locked p), sums in index
 *     order.
 *   Loop: p2 = EM(p1), p3 = EM(p2, filtering); over unlocked entries r =
 *     p2 - p1, v = (p3 - p2) - r, g2 = |r|^2 / |v|^2 (1 when |v|^2 <=
 *     1e-300), g = -sqrt(g2), p3 = p1 - 2 g r + g2 v, or p3 / 10 when that
 *     is <= 1e-300; then (locking) an unlocked entry locks when p3 <
 *     1e-300 or |p3 - p1| < lock_thr p3; p1 = EM(p3); converged when no
 *     entry has p3 >= stop_min and |p3 - p1| / p3 >= stop_rel. */
typedef struct {
	uint32_t T, C, z;
	const uint32_t *off, *ref;
	const double *cnt, *w;
	double thr;
	uint8_t *lock, *alive, *dead;
	double *P0, *bank;
	uint32_t removed_members, removed_classes;
} o_he;

static void he_estep(o_he *S, const double *p, double *acc, int filt)
{
	uint32_t c, k, t;
	const int has_z = S->z < S->T;
	double B = 0;
	for (c = 0; c < S->C; ++c)
		if (S->dead[c])
			B += S->bank[c];
	memset(acc, 0, S->T * sizeof(double));
	if (has_z && !S->lock[S->z] && B >= S->thr)
		acc[S->z] = B;
	for (c = 0; c < S->C; ++c) {
		double total, zpost = 0;
		uint32_t live = 0;
		if (S->dead[c])
			continue;
		total = S->P0[c];
		for (k = S->off[c]; k < S->off[c + 1]; ++k)
			if (S->alive[k])
				total += p[S->ref[k]] * (S->w ? S->w[k] : 1.0);
		if (total < 1e-300) {
			if (filt) {
				S->dead[c] = 1;
				S->bank[c] = S->cnt[c];
				S->removed_classes++;
			}
			continue;
		}
		for (k = S->off[c]; k < S->off[c + 1]; ++k) {
			double post;
			if (!S->alive[k])
				continue;
			t = S->ref[k];
			post = (p[t] * (S->w ? S->w[k] : 1.0)) / total;
			if (t == S->z) {
				zpost = post;
				if (post >= S->thr && !S->lock[t])
					acc[t] += S->cnt[c] * post;
				continue;
			}
			if (post < S->thr) {
				if (filt) {
					S->alive[k] = 0;
					S->removed_members++;
				}
			} else if (S->lock[t]) {
				if (filt) {
					S->alive[k] = 0;
					S->removed_members++;
					S->P0[c] += p[t] *
					             (S->w ? S->w[k] : 1.0);
				}
			} else {
				acc[t] += S->cnt[c] * post;
				++live;
			}
		}
		if (filt && live == 0) {
			S->dead[c] = 1;
			S->bank[c] = S->cnt[c] * zpost;
			S->removed_classes++;
		}
	}
}

static void he_mstep(const o_he *S, const double *p, const double *acc,
                     double *out)
{
	uint32_t t;
	double sum = 0, l_sum = 0, f;
	for (t = 0; t < S->T; ++t) {
		if (S->lock[t])
			l_sum += p[t];
		else
			sum += acc[t];
	}
	f = 1 / sum * (1 - l_sum);
	for (t = 0; t < S->T; ++t)
		out[t] = S->lock[t] ? p[t] : acc[t] * f;
}

int oracle_hera_em(uint32_t n_tx, uint32_t n_cls, const uint32_t *off,
                   const uint32_t *ref, const double *cnt, const double *w,
                   const oracle_he_opts *o, const double *p0, double *p_out,
                   oracle_he_stats *st)
{
	o_he S;
	uint32_t t, k = 0, conv = 0, nz = n_cls ? off[n_cls] : 0;
	double *p1 = malloc(n_tx * sizeof(double));
	double *p2 = malloc(n_tx * sizeof(double));
	double *p3 = malloc(n_tx * sizeof(double));
	double *acc = malloc(n_tx * sizeof(double));
	memset(&S, 0, sizeof(S));
	S.T = n_tx;
	S.C = n_cls;
	S.z = o->noise_tx;
	S.off = off;
	S.ref = ref;
	S.cnt = cnt;
	S.w = w;
	S.thr = o->filter_thr;
	S.lock = calloc(n_tx ? n_tx : 1, 1);
	S.alive = malloc(nz ? nz : 1);
	S.dead = calloc(n_cls ? n_cls : 1, 1);
	S.P0 = calloc(n_cls ? n_cls : 1, sizeof(double));
	S.bank = calloc(n_cls ? n_cls : 1, sizeof(double));
	if (!p1 || !p2 || !p3 || !acc || !S.lock || !S.alive || !S.dead ||
	    !S.P0 || !S.bank || n_tx == 0)
		return 1;
	memset(S.alive, 1, nz ? nz : 1);
	/* trans_uniform_init (transdis.c:32-40): 1 / (trans_cnt + 1) over the
	 * transcripts and the noise */
	for (t = 0; t < n_tx; ++t)
		p1[t] = p0 ? p0[t] : 1.0 / n_tx;
	while (k < o->max_iter) {
		double r2 = 0, v2 = 0, g2, g;
		uint32_t bad = 0;
		he_estep(&S, p1, acc, 0);
		he_mstep(&S, p1, acc, p2);
		he_estep(&S, p2, acc, o->filtering);
		he_mstep(&S, p2, acc, p3);
		for (t = 0; t < n_tx; ++t)
			if (!S.lock[t]) {
				double r = p2[t] - p1[t];
				double v = p3[t] - p2[t] - r;
				r2 += r * r;
				v2 += v * v;
			}
		g2 = v2 <= 1e-300 ? 1 : r2 / v2;
		g = -sqrt(g2);
		for (t = 0; t < n_tx; ++t)
			if (!S.lock[t]) {
				double r = p2[t] - p1[t];
				double v = p3[t] - p2[t] - r;
				double nv = p1[t] - 2 * g * r + g2 * v;
				if (nv <= 1e-300)
					nv = p3[t] / 10;
				p3[t] = nv;
			}
		if (o->locking)
			for (t = 0; t < n_tx; ++t)
				if (!S.lock[t]) {
					double diff = fabs(p3[t] - p1[t]);
					S.lock[t] = p3[t] < 1e-300 ||
					            diff < o->lock_thr * p3[t];
				}
		he_estep(&S, p3, acc, 0);
		he_mstep(&S, p3, acc, p1);
		++k;
		for (t = 0; t < n_tx; ++t) {
			double diff = fabs(p3[t] - p1[t]);
			if (p3[t] >= o->stop_min) {
				diff /= p3[t];
				bad += diff >= o->stop_rel;
			}
		}
		if (bad == 0) {
			conv = 1;
			break;
		}
	}
	memcpy(p_out, p1, n_tx * sizeof(double));
	if (st) {
		st->iters = k;
		st->converged = conv;
		st->n_locked = 0;
		for (t = 0; t < n_tx; ++t)
			st->n_locked += S.lock[t];
		st->removed_members = S.removed_members;
		st->removed_classes = S.removed_classes;
	}
	free(p1);
	free(p2);
	free(p3);
	free(acc);
	free(S.lock);
	free(S.alive);
	free(S.dead);
	free(S.P0);
	free(S.bank);
	return 0;
}
