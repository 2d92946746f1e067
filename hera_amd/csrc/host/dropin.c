/*
 * dropin.c -- hera's estimate_count (src/EM.c:315-387) on the GPU.
 *
 * The class table behind CLASS is flattened into CSR once and kept on the
 * device; it is rebuilt when CLASS, its class array, its size or REF_INF
 * change.  Concurrent callers (hera's bootstrap threads, src/EM.c:243-255)
 * share the handle, which serialises their solves; each holds a reference
 * to it for the length of its call.
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

/* The cached handle and its users.  The cache holds one reference and
 * every estimate_count call in flight holds one more, so a key change or
 * hem_dropin_release while other threads are solving only detaches the
 * handle: the last holder destroys it. */
typedef struct {
	hem_problem *p;
	unsigned refs;
} dropin_ent;

static pthread_mutex_t dropin_mu = PTHREAD_MUTEX_INITIALIZER;
static dropin_ent *dropin_cur;
static const void *dropin_key[4];
static uint32_t dropin_n[2];
static uint32_t bootstrap_cnt;

static void die(const char *what)
{
	fprintf(stderr, "ERROR: %s: %s\n", what, hem_last_error());
	exit(EXIT_FAILURE);
}

/* Drops one reference; dropin_mu held. */
static void ent_put_locked(dropin_ent *e)
{
	if (e && --e->refs == 0) {
		hem_problem_destroy(e->p);
		free(e);
	}
}

static hem_problem *table_upload(void)
{
	uint32_t C = CLASS->n, i;
	uint64_t nnz = 0;
	uint32_t *off = malloc(((size_t)C + 1) * 4);
	double *cnt = malloc((C ? C : 1) * sizeof(double));
	uint32_t *ref;
	hem_classes hc;
	hem_problem *p = NULL;
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
	rc = hem_problem_create(&hc, &p);
	free(off);
	free(ref);
	free(cnt);
	if (rc)
		die("estimate_count: class table upload");
	return p;
}

/* The handle for the current CLASS / REF_INF, with a reference taken for
 * the caller (released by dropin_put). */
static dropin_ent *dropin_get(void)
{
	dropin_ent *e;
	pthread_mutex_lock(&dropin_mu);
	if (dropin_cur && dropin_key[0] == CLASS &&
	    dropin_key[1] == CLASS->cls && dropin_key[2] == REF_INF &&
	    dropin_key[3] == REF_INF->len && dropin_n[0] == CLASS->n &&
	    dropin_n[1] == REF_INF->n) {
		e = dropin_cur;
		++e->refs;
		pthread_mutex_unlock(&dropin_mu);
		return e;
	}
	ent_put_locked(dropin_cur);
	dropin_cur = NULL;
	e = malloc(sizeof(*e));
	if (!e)
		die("estimate_count: host allocation");
	e->p = table_upload();
	e->refs = 2;
	dropin_cur = e;
	dropin_key[0] = CLASS;
	dropin_key[1] = CLASS->cls;
	dropin_key[2] = REF_INF;
	dropin_key[3] = REF_INF->len;
	dropin_n[0] = CLASS->n;
	dropin_n[1] = REF_INF->n;
	pthread_mutex_unlock(&dropin_mu);
	return e;
}

static void dropin_put(dropin_ent *e)
{
	pthread_mutex_lock(&dropin_mu);
	ent_put_locked(e);
	pthread_mutex_unlock(&dropin_mu);
}

EM_val *estimate_count(unsigned int *count, unsigned int order)
{
	hem_opts o;
	hem_stats st;
	dropin_ent *ent;
	EM_val *em_val;
	uint32_t T;

	if (!REF_INF || !CLASS || !count) {
		fprintf(stderr, "ERROR: estimate_count: REF_INF, CLASS or count "
		                "is NULL\n");
		exit(EXIT_FAILURE);
	}
	T = REF_INF->n;
	ent = dropin_get();
	em_val = malloc(sizeof(EM_val));
	if (!em_val)
		die("estimate_count: host allocation");
	em_val->grama = calloc(T, sizeof(double));
	em_val->overlap = calloc(T, sizeof(double));
	if (!em_val->grama || !em_val->overlap)
		die("estimate_count: host allocation");
	hem_opts_default(&o);
	if (hem_solve(ent->p, count, &o, em_val->grama, em_val->overlap, NULL,
	              &st))
		die("estimate_count");
	dropin_put(ent);

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
	ent_put_locked(dropin_cur);
	dropin_cur = NULL;
	memset(dropin_key, 0, sizeof(dropin_key));
	pthread_mutex_unlock(&dropin_mu);
	return HEM_OK;
}
