/*
 * gen.c -- seeded synthetic class tables in hera's layout (hera_em_gen.h).
 *
 * Every random quantity is a Philox draw named by (stage key, index), so the
 * table depends only on the parameters and the seed.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/hera_em.h"
#include "../../../include/hera_em_gen.h"
#include "hem_internal.h"
#include "philox.h"

enum { ST_CLUSTER = 1, ST_LEN = 2, ST_CLASS = 3, ST_PERM = 4 };

void hem_gen_default(hem_gen_params *g, uint32_t n_tx, uint32_t n_cls_all,
                     uint64_t n_frag, uint64_t seed)
{
	memset(g, 0, sizeof(*g));
	g->n_tx = n_tx;
	g->n_cls_all = n_cls_all;
	g->n_frag = n_frag;
	g->cluster_min = 2;
	g->cluster_max = 16;
	g->len_mu = 7.3;
	g->len_sigma = 0.8;
	g->len_min = 200;
	g->len_max = 20000;
	g->geom_p = 0.35;
	g->cls_max = 32;
	g->zipf_s = 1.1;
	g->seed = seed;
}

void hem_gen_free(hem_gen_problem *p)
{
	if (!p)
		return;
	free(p->cls_off);
	free(p->cls_ref);
	free(p->cls_count);
	free(p->tx_len);
	free(p->unique);
	free(p->cluster_off);
	free(p);
}

static int cmp_u32(const void *a, const void *b)
{
	uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
	return x < y ? -1 : x > y;
}

int hem_generate(const hem_gen_params *g, hem_gen_problem **out)
{
	hem_gen_problem *p;
	uint32_t T, Ca, i, n_cl = 0;
	uint32_t *csize = NULL, *cfirst = NULL, *members = NULL, *perm = NULL;
	uint8_t *csz = NULL;
	double W = 0;
	uint64_t total = 0;
	uint64_t kc, kl, kk, kp;

	if (!g || !out || g->n_tx < 2 || g->cluster_min < 1 ||
	    g->cluster_max < g->cluster_min || g->cls_max < 1 ||
	    g->cls_max > 255 || !(g->geom_p > 0 && g->geom_p < 1) ||
	    g->len_min < 1 || g->len_max < g->len_min)
		return hem_fail(HEM_ERR_INVALID, "bad generator parameters");
	T = g->n_tx;
	Ca = g->n_cls_all;
	kc = hem_key(g->seed, ST_CLUSTER);
	kl = hem_key(g->seed, ST_LEN);
	kk = hem_key(g->seed, ST_CLASS);
	kp = hem_key(g->seed, ST_PERM);

	p = calloc(1, sizeof(*p));
	if (!p)
		return hem_fail(HEM_ERR_NOMEM, "host allocation failed");
	p->n_tx = T;
	p->tx_len = malloc((size_t)T * 4);
	p->unique = calloc(T, 4);
	p->cluster_off = malloc(((size_t)T + 1) * 4);
	csz = malloc(Ca ? Ca : 1);
	cfirst = malloc((Ca ? Ca : 1) * (size_t)4);
	csize = NULL;
	perm = malloc((Ca ? Ca : 1) * (size_t)4);
	if (!p->tx_len || !p->unique || !p->cluster_off || !csz || !cfirst ||
	    !perm)
		goto nomem;

	/* 1. gene clusters: uniform sizes, contiguous ids, last one cut */
	p->cluster_off[0] = 0;
	for (uint32_t t = 0; t < T; ++n_cl) {
		hem_u32x4 h = hem_philox(kc, n_cl, 0, 0, 0);
		uint32_t span = g->cluster_max - g->cluster_min + 1;
		uint32_t sz = g->cluster_min +
		              (uint32_t)(hem_u01(h.v[0], h.v[1]) * span);
		if (sz > g->cluster_max)
			sz = g->cluster_max;
		t = (T - t < sz) ? T : t + sz;
		p->cluster_off[n_cl + 1] = t;
	}
	p->n_clusters = n_cl;

	/* 2. integer lengths: lognormal by Box-Muller, clamped, rounded */
	for (i = 0; i < T; ++i) {
		hem_u32x4 h = hem_philox(kl, i, 0, 0, 0);
		double u1 = hem_u01(h.v[0], h.v[1]), u2 = hem_u01(h.v[2], h.v[3]);
		double z = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
		double x = exp(g->len_mu + g->len_sigma * z);
		if (x < g->len_min)
			x = g->len_min;
		if (x > g->len_max)
			x = g->len_max;
		p->tx_len[i] = (uint32_t)llround(x);
	}

	/* 3. class sizes and clusters; members drawn in the fill pass below */
	{
		const double lq = log(1.0 - g->geom_p);
		uint64_t nnz = 0;
		uint32_t nm = 0;
		for (i = 0; i < Ca; ++i) {
			hem_u32x4 h = hem_philox(kk, i, 0, 0, 0);
			uint32_t cl = (uint32_t)(hem_u01(h.v[0], h.v[1]) * n_cl);
			uint32_t m, s;
			double gs;
			if (cl >= n_cl)
				cl = n_cl - 1;
			m = p->cluster_off[cl + 1] - p->cluster_off[cl];
			gs = floor(log(hem_u01(h.v[2], h.v[3])) / lq) + 1.0;
			s = gs > g->cls_max ? g->cls_max : (uint32_t)gs;
			if (s > m)
				s = m;
			csz[i] = (uint8_t)s;
			cfirst[i] = cl;
			if (s >= 2) {
				nnz += s;
				++nm;
			}
		}
		if (nnz > 0xFFFFFFFFull)
			goto nomem;
		p->n_cls = nm;
		p->nnz = nnz;
		p->cls_off = malloc(((size_t)nm + 1) * 4);
		p->cls_ref = malloc((nnz ? nnz : 1) * 4);
		p->cls_count = malloc((nm ? nm : 1) * 8);
		members = malloc(g->cluster_max * 4);
		if (!p->cls_off || !p->cls_ref || !p->cls_count || !members)
			goto nomem;
	}

	/* 4. Zipf ranks: Fisher-Yates permutation, weights r^-s */
	for (i = 0; i < Ca; ++i)
		perm[i] = i;
	for (i = 0; i + 1 < Ca; ++i) {
		hem_u32x4 h = hem_philox(kp, i, 0, 0, 0);
		uint32_t j = i + (uint32_t)(hem_u01(h.v[0], h.v[1]) * (Ca - i));
		uint32_t tmp;
		if (j >= Ca)
			j = Ca - 1;
		tmp = perm[i];
		perm[i] = perm[j];
		perm[j] = tmp;
	}
	for (i = 0; i < Ca; ++i)
		W += pow((double)i + 1.0, -g->zipf_s);

	/* 5. fill: members, counts; singletons fold into unique counts */
	{
		uint32_t c = 0;
		uint64_t z = 0;
		p->cls_off[0] = 0;
		for (i = 0; i < Ca; ++i) {
			uint32_t cl = cfirst[i], s = csz[i];
			uint32_t base = p->cluster_off[cl];
			uint32_t m = p->cluster_off[cl + 1] - base;
			double w = pow((double)perm[i] + 1.0, -g->zipf_s);
			double x = llround((double)g->n_frag * w / W);
			uint64_t cnt = x < 1 ? 1 : (uint64_t)x;
			total += cnt;
			for (uint32_t a = 0; a < m; ++a)
				members[a] = base + a;
			/* partial Fisher-Yates: first s of a random order */
			for (uint32_t a = 0; a < s; ++a) {
				hem_u32x4 h = hem_philox(kk, i, a + 1, 0, 0);
				uint32_t b = a + (uint32_t)(hem_u01(h.v[0], h.v[1]) *
				                            (m - a));
				uint32_t tmp;
				if (b >= m)
					b = m - 1;
				tmp = members[a];
				members[a] = members[b];
				members[b] = tmp;
			}
			if (s < 2) {
				uint64_t u = (uint64_t)p->unique[members[0]] + cnt;
				if (u > 0xFFFFFFFFull)
					goto nomem;
				p->unique[members[0]] = (uint32_t)u;
				continue;
			}
			qsort(members, s, 4, cmp_u32);
			memcpy(p->cls_ref + z, members, (size_t)s * 4);
			z += s;
			p->cls_count[c] = (double)cnt;
			p->cls_off[++c] = (uint32_t)z;
		}
	}
	p->frag_total = total;
	free(csz);
	free(cfirst);
	free(csize);
	free(members);
	free(perm);
	*out = p;
	return HEM_OK;
nomem:
	free(csz);
	free(cfirst);
	free(csize);
	free(members);
	free(perm);
	hem_gen_free(p);
	return hem_fail(HEM_ERR_NOMEM, "generator allocation or range failure");
}
