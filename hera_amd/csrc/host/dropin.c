/*
 * dropin.c -- hera's estimate_count (src/EM.c:315-387) on the GPU.
 *
 * The class table behind CLASS is flattened into CSR once and kept on the
 * device; it is rebuilt when CLASS, its class array, its size or REF_INF
 * change.  Concurrent callers (hera's bootstrap threads, src/EM.c:243-255)
 * are serialised by the problem handle.
 */
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/hera_dropin.h"
#include "../../../include/hera_em.h"
#include "hem_internal.h"

/* hera defines these (src/hash_align.c:35,42); the weak definitions only
 * let the library load on its own. */
__attribute__((weak)) Ref_inf *REF_INF = NULL;
__attribute__((weak)) Gclass *CLASS = NULL;

/* hera's logger (src/log.c:25), called when hera provides it. */
extern void _log(const char *fmt, ...) __attribute__((weak));

static pthread_mutex_t dropin_mu = PTHREAD_MUTEX_INITIALIZER;
static hem_problem *dropin_prob;
static const void *dropin_key[4];
static uint32_t dropin_n[2];
static uint32_t bootstrap_cnt;

static void die(const char *what)
{
	fprintf(stderr, "ERROR: %s: %s\n", what, hem_last_error());
	exit(EXIT_FAILURE);
}

static hem_problem *cached_problem(void)
{
	hem_problem *p;
	pthread_mutex_lock(&dropin_mu);
	if (dropin_prob && dropin_key[0] == CLASS &&
	    dropin_key[1] == CLASS->cls && dropin_key[2] == REF_INF &&
	    dropin_key[3] == REF_INF->len && dropin_n[0] == CLASS->n &&
	    dropin_n[1] == REF_INF->n) {
		p = dropin_prob;
		pthread_mutex_unlock(&dropin_mu);
		return p;
	}
	hem_problem_destroy(dropin_prob);
	dropin_prob = NULL;
	{
		uint32_t C = CLASS->n, i;
		uint64_t nnz = 0;
		uint32_t *off = malloc(((size_t)C + 1) * 4);
		double *cnt = malloc((C ? C : 1) * sizeof(double));
		uint32_t *ref;
		hem_classes hc;
		int rc;
		if (!off || !cnt)
			die("estimate_count: host allocation");
		off[0] = 0;
		for (i = 0; i < C; ++i) {
			nnz += CLASS->cls[i].n;
			if (nnz > 0xFFFFFFFFull) {
				fprintf(stderr, "ERROR: estimate_count: more than "
				                "2^32-1 class members\n");
				exit(EXIT_FAILURE);
			}
			off[i + 1] = (uint32_t)nnz;
			cnt[i] = CLASS->cls[i].count;
		}
		ref = malloc((nnz ? nnz : 1) * 4);
		if (!ref)
			die("estimate_count: host allocation");
		for (i = 0; i < C; ++i)
			memcpy(ref + off[i], CLASS->cls[i].ref,
			       (size_t)CLASS->cls[i].n * 4);
		hc.n_tx = REF_INF->n;
		hc.n_cls = C;
		hc.cls_off = off;
		hc.cls_ref = ref;
		hc.cls_count = cnt;
		hc.tx_len = REF_INF->len;
		hc.member_weight = NULL;
		rc = hem_problem_create(&hc, &dropin_prob);
		free(off);
		free(ref);
		free(cnt);
		if (rc)
			die("estimate_count: class table upload");
	}
	dropin_key[0] = CLASS;
	dropin_key[1] = CLASS->cls;
	dropin_key[2] = REF_INF;
	dropin_key[3] = REF_INF->len;
	dropin_n[0] = CLASS->n;
	dropin_n[1] = REF_INF->n;
	p = dropin_prob;
	pthread_mutex_unlock(&dropin_mu);
	return p;
}

EM_val *estimate_count(unsigned int *count, unsigned int order)
{
	hem_opts o;
	hem_stats st;
	hem_problem *p;
	EM_val *em_val;
	uint32_t T;

	if (!REF_INF || !CLASS || !count) {
		fprintf(stderr, "ERROR: estimate_count: REF_INF, CLASS or count "
		                "is NULL\n");
		exit(EXIT_FAILURE);
	}
	T = REF_INF->n;
	p = cached_problem();
	em_val = malloc(sizeof(EM_val));
	if (!em_val)
		die("estimate_count: host allocation");
	em_val->grama = calloc(T, sizeof(double));
	em_val->overlap = calloc(T, sizeof(double));
	if (!em_val->grama || !em_val->overlap)
		die("estimate_count: host allocation");
	hem_opts_default(&o);
	if (hem_solve(p, count, &o, em_val->grama, em_val->overlap, NULL, &st))
		die("estimate_count");

	/* src/EM.c:375-384 */
	if (order == 0) {
		fprintf(stderr, "Finish EM with %u rounds\n", st.iters);
		if (_log)
			_log("Finish EM with %u rounds\n", st.iters);
	} else {
		uint32_t b = __atomic_add_fetch(&bootstrap_cnt, 1,
		                                __ATOMIC_RELAXED);
		fprintf(stderr, "Finish bootstrap number %u with %u rounds\n", b,
		        st.iters);
		if (_log)
			_log("Finish bootstrap number %u with %u rounds\n", b,
			     st.iters);
	}
	return em_val;
}

__attribute__((weak)) void destroy_emVal(EM_val *em_val)
{
	if (!em_val)
		return;
	free(em_val->overlap);
	free(em_val->grama);
	free(em_val);
}

int hem_dropin_release(void)
{
	pthread_mutex_lock(&dropin_mu);
	hem_problem_destroy(dropin_prob);
	dropin_prob = NULL;
	memset(dropin_key, 0, sizeof(dropin_key));
	pthread_mutex_unlock(&dropin_mu);
	return HEM_OK;
}
