/* hem_internal.h -- host-side state of the C library (not installed). */
#ifndef HEM_INTERNAL_H_
#define HEM_INTERNAL_H_

#include <pthread.h>
#include <stdint.h>

#include "../../../include/hera_em.h"
#include "../dev/hem_dev.h"

/* Iterations per captured graph.  Iterations past convergence are no-ops
 * (every kernel reads the control block's done flag first). */
#define HEM_ITERS_PER_GRAPH 4
/* Distinct option sets whose graphs a handle keeps. */
#define HEM_GRAPH_CACHE 8

struct hem_problem {
	pthread_mutex_t mu;
	uint32_t n_tx, n_cls;
	uint64_t nnz;
	uint64_t dev_bytes;
	int group_cls, group_tx;
	void *stream;
	void *ev0, *ev1;

	/* structure (device) */
	uint32_t *d_cls_off, *d_cls_ref, *d_tx_off, *d_tx_ref, *d_len;
	float *d_cnt32;
	double *d_cnt64;
	double *d_nd;      /* n_cls pairs {n_c, D_c} of the FP64 class pass */
	hemk_prob dp;

	/* one solver instance (device) */
	double *d_A, *d_B, *d_C, *d_D;
	float *d_A32, *d_B32, *d_D32;
	void *d_rho;
	uint32_t *d_unique, *d_flags;
	double *d_part1, *d_part2, *d_out;
	hemk_ctl *d_ctl;
	hemk_ctl *h_ctl; /* pinned */
	uint32_t n_flags;

	/* captured iteration graphs and the options each was captured for;
	 * kept until the handle is destroyed */
	void *graph[HEM_GRAPH_CACHE];
	int graph_key[HEM_GRAPH_CACHE];
	int n_graph;
};

/* kernel kinds for per-kernel timing (hem_kernel_profile) */
enum { HEM_PK_DENOM = 0, HEM_PK_GATHER, HEM_PK_NORM1, HEM_PK_NORM2,
       HEM_PK_EXTRAP, HEM_PK_ITER_END };

/* error.c */
int hem_fail(int code, const char *fmt, ...);
int hem_dev_fail(int hip_code, const char *what);
int hem_device_ready(void);

#define HEM_TRY_DEV(call, what)                                   \
	do {                                                      \
		int e_ = (call);                                  \
		if (e_)                                           \
			return hem_dev_fail(e_, (what));          \
	} while (0)

#endif
